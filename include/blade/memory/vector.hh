// Reference-counted tensor over pinned host, HIP device or managed memory
// (reference include/blade/memory/vector.hh:15-278; same interface: copy
// shares, move steals, the last owner frees).  The count is atomic, so
// handles may be shared across host threads.
#ifndef BLADE_MEMORY_VECTOR_HH
#define BLADE_MEMORY_VECTOR_HH

#include <hip/hip_runtime_api.h>

#include <atomic>
#include <utility>

#include "blade/logger.hh"
#include "blade/macros.hh"
#include "blade/memory/profiler.hh"
#include "blade/memory/types.hh"

namespace Blade {

template <Device DeviceId, typename DataType, typename Shape>
struct Vector {
 public:
    using Type = DataType;
    using ShapeType = Shape;
    static constexpr Device device_id = DeviceId;

    Vector() : _shape(), _data(nullptr), _refs(nullptr), _unified(false) {}

    // Non-owning view of existing memory.
    explicit Vector(void* ptr, const typename Shape::Type& shape, bool unified = false)
        : _shape(shape), _data(static_cast<DataType*>(ptr)), _refs(nullptr), _unified(unified) {}

    explicit Vector(const typename Shape::Type& shape, bool unified = false)
        : _shape(shape), _data(nullptr), _refs(new std::atomic<U64>(1)), _unified(unified) {
        if (Profiler::IsCapturing()) {
            if (_unified) Profiler::RegisterUnifiedAllocation(size_bytes());
            else if (DeviceId == Device::CPU) Profiler::RegisterCpuAllocation(size_bytes());
            else Profiler::RegisterCudaAllocation(size_bytes());
            return;
        }
        if (size_bytes() == 0) return;
        hipError_t e = hipSuccess;
        if constexpr (DeviceId == Device::CPU) {
            e = hipHostMalloc(reinterpret_cast<void**>(&_data), size_bytes(), hipHostMallocDefault);
        } else {
            e = unified ? hipMallocManaged(reinterpret_cast<void**>(&_data), size_bytes())
                        : hipMalloc(reinterpret_cast<void**>(&_data), size_bytes());
        }
        if (e != hipSuccess) {
            BL_FATAL("Failed to allocate {} bytes of {} memory: {}", size_bytes(),
                     DeviceInfo<DeviceId>::name, hipGetErrorString(e));
            delete _refs;
            _refs = nullptr;
            _data = nullptr;
            throw Result::CUDA_ERROR;
        }
    }

    Vector(const Vector& other)
        : _shape(other._shape), _data(other._data), _refs(other._refs), _unified(other._unified) {
        increaseRefCount();
    }

    Vector(Vector&& other) noexcept : _shape(), _data(nullptr), _refs(nullptr), _unified(false) {
        std::swap(_data, other._data);
        std::swap(_refs, other._refs);
        std::swap(_shape, other._shape);
        std::swap(_unified, other._unified);
    }

    Vector& operator=(const Vector& other) {
        if (this == &other) return *this;
        decreaseRefCount();
        _data = other._data;
        _refs = other._refs;
        _shape = other._shape;
        _unified = other._unified;
        increaseRefCount();
        return *this;
    }

    Vector& operator=(Vector&& other) noexcept {
        if (this == &other) return *this;
        decreaseRefCount();
        reset();
        std::swap(_data, other._data);
        std::swap(_refs, other._refs);
        std::swap(_shape, other._shape);
        std::swap(_unified, other._unified);
        return *this;
    }

    ~Vector() { decreaseRefCount(); }

    DataType* data() const noexcept { return _data; }
    U64 refs() const noexcept { return _refs ? _refs->load() : 0; }
    const char* type() const noexcept { return TypeInfo<DataType>::name; }
    const char* device() const noexcept { return DeviceInfo<DeviceId>::name; }
    U64 size() const noexcept { return _shape.size(); }
    U64 size_bytes() const noexcept { return size() * sizeof(DataType); }
    const Shape& shape() const { return _shape; }
    bool unified() const { return _unified; }
    bool empty() const noexcept { return _data == nullptr; }

    DataType& operator[](const typename Shape::Type& c) { return _data[_shape.shapeToOffset(c)]; }
    const DataType& operator[](const typename Shape::Type& c) const { return _data[_shape.shapeToOffset(c)]; }
    DataType& operator[](U64 idx) { return _data[idx]; }
    const DataType& operator[](U64 idx) const { return _data[idx]; }

    DataType* begin() { return _data; }
    DataType* end() { return _data + size(); }
    const DataType* begin() const { return _data; }
    const DataType* end() const { return _data + size(); }

    Result reshape(const Shape& shape) {
        if (shape.size() != _shape.size()) return Result::ERROR;
        _shape = shape;
        return Result::SUCCESS;
    }

 private:
    Shape _shape;
    DataType* _data;
    std::atomic<U64>* _refs;
    bool _unified;

    void increaseRefCount() {
        if (_refs) _refs->fetch_add(1);
    }

    void decreaseRefCount() {
        if (!_refs) return;
        if (_refs->fetch_sub(1) != 1) return;
        delete _refs;
        if (Profiler::IsCapturing()) {
            if (_unified) Profiler::RegisterUnifiedDeallocation(size_bytes());
            else if (DeviceId == Device::CPU) Profiler::RegisterCpuDeallocation(size_bytes());
            else Profiler::RegisterCudaDeallocation(size_bytes());
        } else if (_data) {
            const hipError_t e = (DeviceId == Device::CPU) ? hipHostFree(_data) : hipFree(_data);
            if (e != hipSuccess) BL_FATAL("Failed to deallocate {} memory.", DeviceInfo<DeviceId>::name);
        }
        reset();
    }

    void reset() {
        _data = nullptr;
        _refs = nullptr;
        _shape = Shape();
        _unified = false;
    }
};

}  // namespace Blade

#endif
