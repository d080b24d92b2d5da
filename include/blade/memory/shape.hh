// Row-major N-dimensional shapes (reference include/blade/memory/shape.hh:12-110;
// same row-major strides and offset <-> coordinate mapping).
#ifndef BLADE_MEMORY_SHAPE_HH
#define BLADE_MEMORY_SHAPE_HH

#include <array>
#include <ostream>

#include "blade/memory/types.hh"

namespace Blade {

template <U64 Dimensions>
struct Shape {
 public:
    using Type = std::array<U64, Dimensions>;

    Shape() : _shape{}, _strides{}, _size(0) {}
    Shape(const Type& shape) : _shape(shape) { computeCache(); }
    Shape(const Shape& other) : _shape(other._shape) { computeCache(); }
    Shape& operator=(const Shape& other) {
        _shape = other._shape;
        computeCache();
        return *this;
    }

    U64 size() const { return _size; }
    const U64* data() const { return _shape.data(); }
    static constexpr U64 dimensions() { return Dimensions; }
    const Type& strides() const { return _strides; }

    // Offset of coordinates `shape` with `offset` added along `axis`.
    U64 shapeToOffset(Type shape, U64 axis = 0, U64 offset = 0) const {
        shape[axis] += offset;
        U64 index = 0;
        for (U64 i = 0; i < Dimensions; i++) index += shape[i] * _strides[i];
        return index;
    }

    Type offsetToShape(U64 index, U64 axis = 0, U64 offset = 0) const {
        Type shape{};
        for (U64 i = 0; i < Dimensions; i++) {
            shape[i] = index / _strides[i];
            index -= shape[i] * _strides[i];
        }
        shape[axis] += offset;
        return shape;
    }

    bool operator==(const Shape& other) const { return _shape == other._shape; }
    bool operator!=(const Shape& other) const { return !(_shape == other._shape); }
    operator Type() const { return _shape; }
    U64& operator[](U64 idx) { return _shape[idx]; }
    const U64& operator[](U64 idx) const { return _shape[idx]; }

 protected:
    Type _shape;
    Type _strides;
    U64 _size;

    void computeCache() {
        _size = 1;
        for (U64 i = 0; i < Dimensions; i++) _size *= _shape[i];
        U64 s = 1;
        for (U64 i = Dimensions; i-- > 0;) {
            _strides[i] = s;
            s *= _shape[i];
        }
    }
};

template <U64 D>
std::ostream& operator<<(std::ostream& os, const Shape<D>& s) {
    os << "[";
    for (U64 i = 0; i < D; i++) os << (i ? ", " : "") << s[i];
    return os << "]";
}

}  // namespace Blade

#endif
