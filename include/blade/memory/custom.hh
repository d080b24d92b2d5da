// Typed shapes and tensors of the ATA path (reference include/blade/memory/custom.hh:10-163):
//   ArrayShape [A, F, T, P], PhasorShape [B, A, F, T, P], DelayShape [B, A], VectorShape [N].
#ifndef BLADE_MEMORY_CUSTOM_HH
#define BLADE_MEMORY_CUSTOM_HH

#include "blade/memory/shape.hh"
#include "blade/memory/vector.hh"

namespace Blade {

struct ArrayShape : public Shape<4> {
    inline static const char* name = "ArrayShape";
    using Shape::Shape;
    ArrayShape(const Shape<4>& s) : Shape<4>(s) {}

    const U64& numberOfAspects() const { return (*this)[0]; }
    const U64& numberOfFrequencyChannels() const { return (*this)[1]; }
    const U64& numberOfTimeSamples() const { return (*this)[2]; }
    const U64& numberOfPolarizations() const { return (*this)[3]; }
};

template <Device DeviceId, typename DataType>
using ArrayTensor = Vector<DeviceId, DataType, ArrayShape>;

struct PhasorShape : public Shape<5> {
    inline static const char* name = "PhasorShape";
    using Shape::Shape;
    PhasorShape(const Shape<5>& s) : Shape<5>(s) {}

    const U64& numberOfBeams() const { return (*this)[0]; }
    const U64& numberOfAntennas() const { return (*this)[1]; }
    const U64& numberOfFrequencyChannels() const { return (*this)[2]; }
    const U64& numberOfTimeSamples() const { return (*this)[3]; }
    const U64& numberOfPolarizations() const { return (*this)[4]; }
};

template <Device DeviceId, typename DataType>
using PhasorTensor = Vector<DeviceId, DataType, PhasorShape>;

struct DelayShape : public Shape<2> {
    inline static const char* name = "DelayShape";
    using Shape::Shape;
    DelayShape(const Shape<2>& s) : Shape<2>(s) {}

    const U64& numberOfBeams() const { return (*this)[0]; }
    const U64& numberOfAntennas() const { return (*this)[1]; }
};

template <Device DeviceId, typename DataType>
using DelayTensor = Vector<DeviceId, DataType, DelayShape>;

struct VectorShape : public Shape<1> {
    inline static const char* name = "VectorShape";
    using Shape::Shape;
    VectorShape(const Shape<1>& s) : Shape<1>(s) {}
};

template <Device DeviceId, typename DataType>
using Tensor = Vector<DeviceId, DataType, VectorShape>;

}  // namespace Blade

#endif
