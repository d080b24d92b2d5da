// Python bindings of the Blade C++ API (pybind11).  One extension module,
// `blade_amd._blade_ext`, whose submodules mirror the reference's nanobind
// extensions one for one (python/blade/ext_*.cc): _const_impl, _hidden_impl,
// _mem_impl, _copy_impl, _pipeline_impl, _runner_impl and one _<name>_impl per
// module with the path [tel_<telescope>.][in_<it>.][out_<ot>.]mod and the
// reference's kwarg names.  The package registers each submodule under its
// reference name (blade_amd._beamformer_impl, ...).
#include <pybind11/functional.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cstring>
#include <sstream>

#include "blade/base.hh"

namespace py = pybind11;
using namespace pybind11::literals;
using namespace Blade;

namespace {

// Pinned or pageable host memory (not device or managed memory).
bool is_host_memory(const void* p) {
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, p) != hipSuccess) {
        (void)hipGetLastError();   // unknown to HIP: pageable host memory
        return true;
    }
    return !attr.isManaged && (attr.type == hipMemoryTypeHost || attr.type == hipMemoryTypeUnregistered);
}

// Result codes surface as exceptions only from constructors (like the
// reference, python/blade/ext_hidden.cc:17-24).
void translate_result() {
    py::register_exception_translator([](std::exception_ptr p) {
        try {
            if (p) std::rethrow_exception(p);
        } catch (const Result& r) {
            PyErr_SetString(PyExc_RuntimeError,
                            ("Blade error (result " + std::to_string(static_cast<int>(r)) + ")").c_str());
        }
    });
}

template <typename T>
struct NumpyOf { using type = T; static constexpr int comps = 1; };
template <> struct NumpyOf<F16> { using type = uint16_t; static constexpr int comps = 1; };
template <> struct NumpyOf<CI8> { using type = int8_t; static constexpr int comps = 2; };
template <> struct NumpyOf<CF16> { using type = uint16_t; static constexpr int comps = 2; };

template <typename T>
py::dtype numpy_dtype() {
    if constexpr (std::is_same<T, F16>::value || std::is_same<T, CF16>::value) return py::dtype("float16");
    else if constexpr (std::is_same<T, CI8>::value) return py::dtype("int8");
    else return py::dtype::of<T>();
}

// Bounds-checked indexing: Python's sequence protocol (list(shape),
// iteration) stops at IndexError.
template <typename Shape>
U64 shape_at(const Shape& s, U64 i) {
    if (i >= Shape::dimensions()) throw py::index_error("shape index out of range");
    return s[i];
}

template <typename Shape>
std::string shape_str(const Shape& s) {
    std::ostringstream os;
    os << s;
    return os.str();
}

// ---------------------------------------------------------------- tensors

template <Device D, typename T, typename Shape>
void bind_vector(py::module_& m, const char* name) {
    using V = Vector<D, T, Shape>;
    py::class_<V> cls(m, name);
    cls.def(py::init<>())
        .def(py::init([](const typename Shape::Type& shape, bool unified) { return V(shape, unified); }),
             "shape"_a, "unified"_a = false)
        .def(py::init([](const Shape& shape, bool unified) { return V(typename Shape::Type(shape), unified); }),
             "shape"_a, "unified"_a = false)
        .def("__repr__", [](V& v) {
            return Log::format("Vector({}, dtype={}, device={}, unified={}, hash={})",
                               shape_str(v.shape()), TypeInfo<T>::name, DeviceInfo<D>::name,
                               v.unified(), reinterpret_cast<uintptr_t>(v.data()));
        })
        .def_property_readonly("device", [](V& v) { return std::string(DeviceInfo<D>::name); })
        .def_property_readonly("unified", [](V& v) { return v.unified(); })
        .def_property_readonly("hash", [](V& v) { return reinterpret_cast<uintptr_t>(v.data()); })
        .def_property_readonly("shape", [](V& v) { return v.shape(); })
        .def("data_ptr", [](V& v) { return reinterpret_cast<uintptr_t>(v.data()); })
        .def("size", [](V& v) { return v.size(); })
        .def("size_bytes", [](V& v) { return v.size_bytes(); });
    if constexpr (D == Device::CPU) {
        // Zero-copy numpy view of pinned host memory.  CI8 and CF16 get a
        // trailing axis of 2 (the reference has no view for them, ext_mem.cc:60-62).
        cls.def("as_numpy", [](V& v) {
            std::vector<py::ssize_t> shape(v.shape().data(), v.shape().data() + Shape::dimensions());
            py::dtype dt = numpy_dtype<T>();
            if (NumpyOf<T>::comps == 2) shape.push_back(2);
            std::vector<py::ssize_t> strides(shape.size());
            py::ssize_t st = dt.itemsize();
            for (size_t k = shape.size(); k-- > 0;) {
                strides[k] = st;
                st *= shape[k];
            }
            auto* keep = new V(v);
            py::capsule owner(keep, [](void* p) { delete static_cast<V*>(p); });
            return py::array(dt, shape, strides, v.data(), owner);
        });
    }
    using DV = Duet<V>;
    py::class_<DV>(m, (std::string(name) + "_duet").c_str())
        .def(py::init([](const typename Shape::Type& shape, bool unified) {
                 return new DV(shape, unified);
             }),
             "shape"_a, "unified"_a = false)
        .def("set", &DV::set)
        .def("at", &DV::at, py::return_value_policy::reference_internal)
        .def("__getitem__", &DV::operator[], py::return_value_policy::reference_internal)
        .def("__call__", [](DV& d) -> V& { return d.fixed(); }, py::return_value_policy::reference_internal);
}

template <Device D, typename T>
void bind_device_type(py::module_& m, const char* name) {
    auto mm = m.def_submodule(name);
    bind_vector<D, T, ArrayShape>(mm, "array_tensor");
    bind_vector<D, T, PhasorShape>(mm, "phasor_tensor");
    bind_vector<D, T, DelayShape>(mm, "delay_tensor");
    bind_vector<D, T, VectorShape>(mm, "tensor");
}

template <Device D>
void bind_device(py::module_& m, const char* name) {
    auto mm = m.def_submodule(name);
    bind_device_type<D, I8>(mm, "i8");
    bind_device_type<D, F16>(mm, "f16");
    bind_device_type<D, F32>(mm, "f32");
    bind_device_type<D, F64>(mm, "f64");
    bind_device_type<D, I32>(mm, "i32");     // output of the F32 -> I32 cast
    bind_device_type<D, CI8>(mm, "ci8");
    bind_device_type<D, CF16>(mm, "cf16");
    bind_device_type<D, CF32>(mm, "cf32");
    bind_device_type<D, CF64>(mm, "cf64");
}

// -------------------------------------------------------------- modules

template <typename M>
void bind_process(py::class_<M, Module, std::shared_ptr<M>>& cls) {
    cls.def("process", [](M& m, U64 counter) { return m.process(counter); }, "counter"_a = 0)
        .def("get_config", &M::getConfig, py::return_value_policy::reference_internal);
}

using ModulePtr = std::shared_ptr<Module>;

template <typename IT, typename OT>
void bind_beamformer(py::module_& m, const char* in_name, const char* out_name) {
    using G = Modules::Beamformer::Generic<IT, OT>;
    using A = Modules::Beamformer::ATA<IT, OT>;
    auto gm = m.def_submodule(in_name).def_submodule(out_name);
    py::class_<G, Module, std::shared_ptr<G>> gen(gm, "_generic_mod");
    py::class_<typename G::Config>(gen, "config")
        .def(py::init([](BOOL inc, BOOL sq, U64 bs, BOOL split) {
                 return typename G::Config{inc, sq, bs, split};
             }),
             "enable_incoherent_beam"_a = false, "enable_incoherent_beam_sqrt"_a = false,
             "block_size"_a = 512, "enable_split_fp16"_a = false);
    py::class_<typename G::Input>(gen, "input")
        .def(py::init([](const ArrayTensor<Device::CUDA, IT>& buf, const PhasorTensor<Device::CUDA, CF32>& ph) {
                 return typename G::Input{buf, ph};
             }),
             "buffer"_a, "phasors"_a, py::keep_alive<0, 1>(), py::keep_alive<0, 2>());
    bind_process<G>(gen);
    gen.def("get_input_buffer", &G::getInputBuffer, py::return_value_policy::reference_internal)
        .def("get_input_phasors", &G::getInputPhasors, py::return_value_policy::reference_internal)
        .def("get_output", &G::getOutputBuffer, py::return_value_policy::reference_internal)
        .def("__repr__", [](G&) { return std::string("Beamformer()"); });
    auto tm = m.def_submodule("tel_ata").def_submodule(in_name).def_submodule(out_name);
    py::class_<A, G, std::shared_ptr<A>>(tm, "mod")
        .def(py::init<const typename G::Config&, const typename G::Input&, const Stream&>(), "config"_a,
             "input"_a, "stream"_a, py::keep_alive<1, 3>());
    tm.attr("config") = gen.attr("config");
    tm.attr("input") = gen.attr("input");
}

void bind_batched_beamformer(py::module_& m) {
    using M = Modules::Beamformer::Batched<CI8, CF32>;
    auto mm = m.def_submodule("in_ci8").def_submodule("out_cf32");
    py::class_<M, Module, std::shared_ptr<M>> cls(mm, "mod");
    py::class_<typename M::Config>(mm, "config")
        .def(py::init([](BOOL split, int nt, int bt, int cpi, int wpc) {
                 return typename M::Config{split, blade_batch_tuning{nt, bt, cpi, wpc}};
             }),
             "enable_split_fp16"_a = false, "time_samples_per_lane"_a = 0, "beam_tiles"_a = 0,
             "chunks_per_item"_a = 0, "workgroups_per_cu"_a = 0);
    py::class_<typename M::Input>(mm, "input")
        .def(py::init([](const std::vector<ArrayTensor<Device::CUDA, CI8>>& bufs,
                         const std::vector<PhasorTensor<Device::CUDA, CF32>>& phasors) {
                 return typename M::Input{bufs, phasors};
             }),
             "buffers"_a, "phasors"_a);
    cls.def(py::init<const typename M::Config&, const typename M::Input&, const Stream&>(), "config"_a,
            "input"_a, "stream"_a)
        .def("process", [](M& m, U64 counter) { return m.process(counter); }, "counter"_a = 0)
        .def("get_config", &M::getConfig, py::return_value_policy::reference_internal)
        .def("get_output", [](M& m) { return m.getOutputBuffers(); })
        .def("__repr__", [](M&) { return std::string("BatchedBeamformer()"); });
}

template <typename IT, typename OT>
void bind_channelizer(py::module_& m, const char* in_name, const char* out_name) {
    using M = Modules::Channelizer<IT, OT>;
    auto mm = m.def_submodule(in_name).def_submodule(out_name);
    py::class_<M, Module, std::shared_ptr<M>> cls(mm, "mod");
    py::class_<typename M::Config>(mm, "config")
        .def(py::init([](U64 rate, U64 bs) { return typename M::Config{rate, bs}; }), "rate"_a,
             "block_size"_a = 512);
    py::class_<typename M::Input>(mm, "input")
        .def(py::init([](const ArrayTensor<Device::CUDA, IT>& buf) { return typename M::Input{buf}; }),
             "buf"_a, py::keep_alive<0, 1>());
    cls.def(py::init<const typename M::Config&, const typename M::Input&, const Stream&>(), "config"_a,
            "input"_a, "stream"_a, py::keep_alive<1, 3>());
    bind_process<M>(cls);
    cls.def("get_input", &M::getInputBuffer, py::return_value_policy::reference_internal)
        .def("get_output", &M::getOutputBuffer, py::return_value_policy::reference_internal);
}

template <typename IT, typename OT>
void bind_polarizer(py::module_& m, const char* in_name, const char* out_name) {
    using M = Modules::Polarizer<IT, OT>;
    auto mm = m.def_submodule(in_name).def_submodule(out_name);
    py::class_<M, Module, std::shared_ptr<M>> cls(mm, "mod");
    py::class_<typename M::Config>(mm, "config")
        .def(py::init([](POL in, POL out, U64 bs) { return typename M::Config{in, out, bs}; }),
             "inputPolarization"_a = POL::XY, "outputPolarization"_a = POL::LR, "block_size"_a = 512);
    py::class_<typename M::Input>(mm, "input")
        .def(py::init([](const ArrayTensor<Device::CUDA, IT>& buf) { return typename M::Input{buf}; }),
             "buf"_a, py::keep_alive<0, 1>());
    cls.def(py::init<const typename M::Config&, const typename M::Input&, const Stream&>(), "config"_a,
            "input"_a, "stream"_a, py::keep_alive<1, 3>());
    bind_process<M>(cls);
    cls.def("get_input", &M::getInputBuffer, py::return_value_policy::reference_internal)
        .def("get_output", &M::getOutputBuffer, py::return_value_policy::reference_internal);
}

template <typename IT, typename OT>
void bind_detector(py::module_& m, const char* in_name, const char* out_name) {
    using M = Modules::Detector<IT, OT>;
    auto mm = m.def_submodule(in_name).def_submodule(out_name);
    py::class_<M, Module, std::shared_ptr<M>> cls(mm, "mod");
    py::class_<typename M::Config>(mm, "config")
        .def(py::init([](U64 intg, U64 npol, U64 bs) { return typename M::Config{intg, npol, bs}; }),
             "integration_size"_a, "number_of_output_polarizations"_a = 4, "block_size"_a = 512);
    py::class_<typename M::Input>(mm, "input")
        .def(py::init([](const ArrayTensor<Device::CUDA, IT>& buf) { return typename M::Input{buf}; }),
             "buf"_a, py::keep_alive<0, 1>());
    cls.def(py::init<const typename M::Config&, const typename M::Input&, const Stream&>(), "config"_a,
            "input"_a, "stream"_a, py::keep_alive<1, 3>());
    bind_process<M>(cls);
    cls.def("get_input", &M::getInputBuffer, py::return_value_policy::reference_internal)
        .def("get_output", &M::getOutputBuffer, py::return_value_policy::reference_internal);
}

template <typename IT, typename OT>
void bind_cast(py::module_& m, const char* in_name, const char* out_name) {
    using M = Modules::Cast<IT, OT>;
    auto mm = m.def_submodule(in_name).def_submodule(out_name);
    py::class_<M, Module, std::shared_ptr<M>> cls(mm, "mod");
    py::class_<typename M::Config>(mm, "config")
        .def(py::init([](U64 bs) { return typename M::Config{bs}; }), "block_size"_a = 512);
    py::class_<typename M::Input>(mm, "input")
        .def(py::init([](const ArrayTensor<Device::CUDA, IT>& buf) { return typename M::Input{buf}; }),
             "buf"_a, py::keep_alive<0, 1>());
    cls.def(py::init<const typename M::Config&, const typename M::Input&, const Stream&>(), "config"_a,
            "input"_a, "stream"_a, py::keep_alive<1, 3>());
    bind_process<M>(cls);
    cls.def("get_input", &M::getInputBuffer, py::return_value_policy::reference_internal)
        .def("get_output", &M::getOutputBuffer, py::return_value_policy::reference_internal);
}

template <typename IT, typename OT>
void bind_gather(py::module_& m, const char* in_name, const char* out_name) {
    using M = Modules::Gather<IT, OT>;
    auto mm = m.def_submodule(in_name).def_submodule(out_name);
    py::class_<M, Module, std::shared_ptr<M>> cls(mm, "mod");
    py::class_<typename M::Config>(mm, "config")
        .def(py::init([](U64 axis, U64 mult, U64 thr, U64 bs) { return typename M::Config{axis, mult, thr, bs}; }),
             "axis"_a, "multiplier"_a, "copy_size_threshold"_a = 512, "block_size"_a = 512);
    py::class_<typename M::Input>(mm, "input")
        .def(py::init([](const ArrayTensor<Device::CUDA, IT>& buf) { return typename M::Input{buf}; }),
             "buf"_a, py::keep_alive<0, 1>());
    cls.def(py::init<const typename M::Config&, const typename M::Input&, const Stream&>(), "config"_a,
            "input"_a, "stream"_a, py::keep_alive<1, 3>());
    bind_process<M>(cls);
    cls.def("get_input", &M::getInputBuffer, py::return_value_policy::reference_internal)
        .def("get_output", &M::getOutputBuffer, py::return_value_policy::reference_internal);
}

template <typename IT, typename OT>
void bind_duplicate(py::module_& m, const char* in_name, const char* out_name) {
    using M = Modules::Duplicate<IT, OT>;
    auto mm = m.def_submodule(in_name).def_submodule(out_name);
    py::class_<M, Module, std::shared_ptr<M>> cls(mm, "mod");
    py::class_<typename M::Config>(mm, "config").def(py::init([]() { return typename M::Config{}; }));
    py::class_<typename M::Input>(mm, "input")
        .def(py::init([](const ArrayTensor<Device::CUDA, IT>& buf) { return typename M::Input{buf}; }),
             "buf"_a, py::keep_alive<0, 1>());
    cls.def(py::init<const typename M::Config&, const typename M::Input&, const Stream&>(), "config"_a,
            "input"_a, "stream"_a, py::keep_alive<1, 3>());
    bind_process<M>(cls);
    cls.def("get_input", &M::getInputBuffer, py::return_value_policy::reference_internal)
        .def("get_output", &M::getOutputBuffer, py::return_value_policy::reference_internal);
}

// GUPPI reader/writer (reference python/blade/ext_guppi.cc:11-100): the
// reference's class names taint_reader.type_ci8 / taint_writer.type_cf32
// and, for bl.module, taint_reader.out_ci8.mod / taint_writer.in_cf32.mod.
template <typename OT>
void bind_guppi_reader(py::module_& taint, const char* out_name, const char* type_name) {
    using M = Modules::Guppi::Reader<OT>;
    auto mm = taint.def_submodule(out_name);
    py::class_<M, Module, std::shared_ptr<M>> cls(mm, "mod");
    py::class_<typename M::Config>(mm, "config")
        .def(py::init([](const std::string& path, U64 nt, U64 nf, U64 na, U64 bs) {
                 return typename M::Config{path, nt, nf, na, bs};
             }),
             "filepath"_a, "step_number_of_time_samples"_a, "step_number_of_frequency_channels"_a,
             "step_number_of_aspects"_a, "block_size"_a = 512);
    py::class_<typename M::Input>(mm, "input").def(py::init([]() { return typename M::Input{}; }));
    cls.def(py::init<const typename M::Config&, const typename M::Input&, const Stream&>(), "config"_a,
            "input"_a, "stream"_a)
        .def(py::init([](const typename M::Config& c, const typename M::Input& i) {
                 return std::make_shared<M>(c, i);
             }),
             "config"_a, "input"_a)
        .def("process", [](M& m, U64 counter) { return m.process(counter); }, "counter"_a = 0)
        .def("get_config", &M::getConfig, py::return_value_policy::reference_internal)
        .def("get_step_output_buffer", &M::getStepOutputBuffer, py::return_value_policy::reference_internal)
        .def("get_step_output_julian_date", &M::getStepOutputJulianDate, py::return_value_policy::reference_internal)
        .def("get_step_output_dut", &M::getStepOutputDut1, py::return_value_policy::reference_internal)
        .def("get_total_output_buffer_shape", &M::getTotalOutputBufferShape)
        .def("get_step_output_buffer_shape", &M::getStepOutputBufferShape)
        .def("get_number_of_steps", &M::getNumberOfSteps)
        .def("get_total_bandwidth", &M::getTotalBandwidth)
        .def("get_channel_bandwidth", &M::getChannelBandwidth)
        .def("get_channel_start_index", &M::getChannelStartIndex)
        .def("get_observation_frequency", &M::getObservationFrequency)
        .def("__repr__", [](const M&) { return std::string("GuppiReader()"); });
    taint.attr(type_name) = cls;
}

template <typename IT>
void bind_guppi_writer(py::module_& taint, const char* in_name, const char* type_name) {
    using M = Modules::Guppi::Writer<IT>;
    auto mm = taint.def_submodule(in_name);
    py::class_<M, Module, std::shared_ptr<M>> cls(mm, "mod");
    py::class_<typename M::Config>(mm, "config")
        .def(py::init([](const std::string& path, bool dio, U64 batches, U64 bs) {
                 return typename M::Config{path, dio, batches, bs};
             }),
             "filepath"_a, "directio"_a, "input_frequency_batches"_a, "block_size"_a = 512);
    py::class_<typename M::Input>(mm, "input")
        .def(py::init([](const ArrayTensor<Device::CPU, IT>& buf) { return typename M::Input{buf}; }),
             "buffer"_a, py::keep_alive<0, 1>());
    cls.def(py::init<const typename M::Config&, const typename M::Input&, const Stream&>(), "config"_a,
            "input"_a, "stream"_a, py::keep_alive<1, 3>())
        .def(py::init([](const typename M::Config& c, const typename M::Input& i) {
                 return std::make_shared<M>(c, i);
             }),
             "config"_a, "input"_a, py::keep_alive<1, 3>())
        .def("process", [](M& m, U64 counter) { return m.process(counter); }, "counter"_a = 0)
        .def("get_config", &M::getConfig, py::return_value_policy::reference_internal)
        .def("get_input", &M::getInputBuffer, py::return_value_policy::reference_internal)
        .def("header_put", [](M& m, const std::string& k, F64 v) { m.headerPut(k, v); })
        .def("header_put", [](M& m, const std::string& k, I64 v) { m.headerPut(k, v); })
        .def("header_put", [](M& m, const std::string& k, const std::string& v) { m.headerPut(k, v); })
        .def("__repr__", [](const M&) { return std::string("GuppiWriter()"); });
    taint.attr(type_name) = cls;
}

template <typename IT, typename OT>
void bind_permutation(py::module_& m, const char* in_name, const char* out_name) {
    using M = Modules::Permutation<IT, OT>;
    auto mm = m.def_submodule(in_name).def_submodule(out_name);
    py::class_<M, Module, std::shared_ptr<M>> cls(mm, "mod");
    py::class_<typename M::Config>(mm, "config")
        .def(py::init([](const std::array<U64, 4>& idx, U64 bs) {
                 return typename M::Config{ArrayShape(idx), bs};
             }),
             "indexes"_a, "block_size"_a = 512);
    py::class_<typename M::Input>(mm, "input")
        .def(py::init([](const ArrayTensor<Device::CUDA, IT>& buf) { return typename M::Input{buf}; }),
             "buf"_a, py::keep_alive<0, 1>());
    cls.def(py::init<const typename M::Config&, const typename M::Input&, const Stream&>(), "config"_a,
            "input"_a, "stream"_a, py::keep_alive<1, 3>());
    bind_process<M>(cls);
    cls.def("get_input", &M::getInputBuffer, py::return_value_policy::reference_internal)
        .def("get_output", &M::getOutputBuffer, py::return_value_policy::reference_internal);
}

template <typename OT>
void bind_phasor(py::module_& m, const char* out_name) {
    using G = Modules::Phasor::Generic<OT>;
    using A = Modules::Phasor::ATA<OT>;
    auto gm = m.def_submodule(out_name);
    py::class_<G, Module, std::shared_ptr<G>> gen(gm, "_generic_mod");
    py::class_<typename G::Config>(gen, "config")
        .def(py::init([](U64 na, U64 nf, U64 np, F64 fobs, F64 cbw, F64 tbw, U64 fstart, U64 refant,
                         LLA ref, RA_DEC bore, std::vector<XYZ> pos,
                         const ArrayTensor<Device::CPU, CF64>& cal, std::vector<RA_DEC> beams, U64 bs) {
                 return typename G::Config{na, nf, np, fobs, cbw, tbw, fstart, refant, ref, bore, pos, cal, beams, bs};
             }),
             "number_of_antennas"_a, "number_of_frequency_channels"_a, "number_of_polarizations"_a,
             "observation_frequency_hz"_a, "channel_bandwidth_hz"_a, "total_bandwidth_hz"_a,
             "frequency_start_index"_a, "reference_antenna_index"_a, "array_reference_position"_a,
             "boresight_coordinate"_a, "antenna_positions"_a, "antenna_calibrations"_a,
             "beam_coordinates"_a, "block_size"_a = 512);
    py::class_<typename G::Input>(gen, "input")
        .def(py::init([](const Tensor<Device::CPU, F64>& jd, const Tensor<Device::CPU, F64>& dut) {
                 return typename G::Input{jd, dut};
             }),
             "block_julian_date"_a, "block_dut1"_a, py::keep_alive<0, 1>(), py::keep_alive<0, 2>());
    bind_process<G>(gen);
    gen.def("get_delays", &G::getOutputDelays, py::return_value_policy::reference_internal)
        .def("get_phasors", &G::getOutputPhasors, py::return_value_policy::reference_internal)
        .def("get_output", &G::getOutputPhasors, py::return_value_policy::reference_internal);
    auto tm = m.def_submodule("tel_ata").def_submodule(out_name);
    py::class_<A, G, std::shared_ptr<A>>(tm, "mod")
        .def(py::init<const typename G::Config&, const typename G::Input&, const Stream&>(), "config"_a,
             "input"_a, "stream"_a, py::keep_alive<1, 3>());
    tm.attr("config") = gen.attr("config");
    tm.attr("input") = gen.attr("input");
}

template <typename IT, typename OT>
void bind_modeb(py::module_& m, const char* in_name, const char* out_name) {
    using M = Bundles::ATA::ModeB<IT, OT>;
    auto mm = m.def_submodule("tel_ata").def_submodule(in_name).def_submodule(out_name);
    py::class_<M, Bundle, std::shared_ptr<M>> cls(mm, "mod");
    py::class_<typename M::Config>(mm, "config")
        .def(py::init([](const std::array<U64, 4>& in_shape, const std::array<U64, 4>& out_shape,
                         U64 rate, BOOL circ, F64 fobs, F64 cbw, F64 tbw, U64 fstart, U64 refant,
                         LLA ref, RA_DEC bore, std::vector<XYZ> pos,
                         const ArrayTensor<Device::CPU, CF64>& cal, std::vector<RA_DEC> beams,
                         BOOL inc, BOOL det, U64 intg, U64 npol, U64 cbs, U64 chbs, U64 pbs, U64 bbs,
                         U64 polbs, U64 dbs, BOOL split) {
                 typename M::Config c{ArrayShape(in_shape), ArrayShape(out_shape), rate, circ, fobs, cbw,
                                      tbw, fstart, refant, ref, bore, pos, cal, beams, inc, det, intg,
                                      npol, cbs, chbs, pbs, bbs, polbs, dbs, split};
                 return c;
             }),
             "input_shape"_a, "output_shape"_a, "pre_beamformer_channelizer_rate"_a,
             "pre_beamformer_polarizer_convert_to_circular"_a = false, "phasor_observation_frequency_hz"_a,
             "phasor_channel_bandwidth_hz"_a, "phasor_total_bandwidth_hz"_a,
             "phasor_frequency_start_index"_a, "phasor_reference_antenna_index"_a,
             "phasor_array_reference_position"_a, "phasor_boresight_coordinate"_a,
             "phasor_antenna_positions"_a, "phasor_antenna_calibrations"_a, "phasor_beam_coordinates"_a,
             "beamformer_incoherent_beam"_a = false, "detector_enable"_a = false,
             "detector_integration_size"_a = 1, "detector_number_of_output_polarizations"_a = 1,
             "cast_block_size"_a = 512, "channelizer_block_size"_a = 512, "phasor_block_size"_a = 512,
             "beamformer_block_size"_a = 512, "polarizer_block_size"_a = 512,
             "detector_block_size"_a = 512, "beamformer_split_fp16"_a = false);
    py::class_<typename M::Input>(mm, "input")
        .def(py::init([](const Tensor<Device::CPU, F64>& dut, const Tensor<Device::CPU, F64>& jd,
                         const ArrayTensor<Device::CUDA, IT>& buf) { return typename M::Input{dut, jd, buf}; }),
             "dut"_a, "julian_date"_a, "buffer"_a, py::keep_alive<0, 1>(), py::keep_alive<0, 2>(),
             py::keep_alive<0, 3>());
    cls.def(py::init<const typename M::Config&, const typename M::Input&, const Stream&>(), "config"_a,
            "input"_a, "stream"_a, py::keep_alive<1, 3>())
        .def("get_output", &M::getOutputBuffer, py::return_value_policy::reference_internal)
        .def("get_config", &M::getConfig, py::return_value_policy::reference_internal);
}

template <typename IT, typename OT>
void bind_modeh(py::module_& m, const char* in_name, const char* out_name) {
    using M = Bundles::Generic::ModeH<IT, OT>;
    auto mm = m.def_submodule(in_name).def_submodule(out_name);
    py::class_<M, Bundle, std::shared_ptr<M>> cls(mm, "mod");
    py::class_<typename M::Config>(mm, "config")
        .def(py::init([](const std::array<U64, 4>& in_shape, const std::array<U64, 4>& out_shape, BOOL circ,
                         U64 intg, U64 npol, U64 cbs, U64 pbs, U64 chbs, U64 dbs) {
                 return typename M::Config{ArrayShape(in_shape), ArrayShape(out_shape), circ, intg, npol,
                                           cbs, pbs, chbs, dbs};
             }),
             "input_shape"_a, "output_shape"_a, "polarizer_convert_to_circular"_a = false,
             "detector_integration_size"_a, "detector_number_of_output_polarizations"_a,
             "cast_block_size"_a = 512, "polarizer_block_size"_a = 512, "channelizer_block_size"_a = 512,
             "detector_block_size"_a = 512);
    py::class_<typename M::Input>(mm, "input")
        .def(py::init([](const ArrayTensor<Device::CUDA, IT>& buf) { return typename M::Input{buf}; }),
             "buffer"_a, py::keep_alive<0, 1>());
    cls.def(py::init<const typename M::Config&, const typename M::Input&, const Stream&>(), "config"_a,
            "input"_a, "stream"_a, py::keep_alive<1, 3>())
        .def("get_output", &M::getOutputBuffer, py::return_value_policy::reference_internal)
        .def("get_config", &M::getConfig, py::return_value_policy::reference_internal);
}

template <typename P>
void bind_pipeline_methods(py::class_<P>& cls) {
    cls.def("compute_complete", &P::computeComplete)
        .def("compute_current_step_count", &P::computeCurrentStepCount)
        .def("compute_steps_per_cycle", &P::computeStepsPerCycle)
        .def("compute_lifetime_cycles", &P::computeLifetimeCycles)
        .def("commited", &P::commited)
        .def("will_output", &P::willOutput)
        .def("compute", &P::compute, "index"_a = 0)
        .def("synchronize", &P::synchronize, "index"_a = 0)
        .def("is_synchronized", &P::isSynchronized, "index"_a = 0)
        .def("enable_graph_capture", &P::enableGraphCapture, "enable"_a = true)
        .def("stream", &P::stream, "index"_a = 0, py::return_value_policy::reference_internal)
        .def("stream_handle", [](P& p, U64 i) { return reinterpret_cast<uintptr_t>(p.stream(i).handle); },
             "index"_a = 0)
        .def("number_of_streams", &P::numberOfStreams)
        .def("add_module", [](P& p, py::object mod) {
            p.addModule(mod.cast<std::shared_ptr<Module>>());
            // Keep the Python wrapper (and the tensors it keeps alive) with the pipeline.
            py::object self = py::cast(&p);
            if (!py::hasattr(self, "_keep")) self.attr("_keep") = py::list();
            self.attr("_keep").cast<py::list>().append(mod);
        });
}

}  // namespace

PYBIND11_MODULE(_blade_ext, m) {
    translate_result();

    // ---------------------------------------------------------- _const_impl
    auto c = m.def_submodule("_const_impl");
    py::enum_<Result>(c, "result")
        .value("success", Result::SUCCESS)
        .value("error", Result::ERROR)
        .value("cuda_error", Result::CUDA_ERROR)
        .value("assertion_error", Result::ASSERTION_ERROR)
        .value("runner_queue_full", Result::RUNNER_QUEUE_FULL)
        .value("runner_queue_empty", Result::RUNNER_QUEUE_EMPTY)
        .value("runner_queue_none_available", Result::RUNNER_QUEUE_NONE_AVAILABLE)
        .value("pipeline_exhausted", Result::PIPELINE_EXHAUSTED);
    py::enum_<POL>(c, "pol")
        .value("x", POL::X).value("y", POL::Y).value("l", POL::L).value("r", POL::R)
        .value("xy", POL::XY).value("lr", POL::LR);
    py::class_<Stream>(c, "stream")
        .def(py::init([]() { return Stream{}; }))
        .def("__repr__", [](Stream& s) { return Log::format("Stream(handle={})", reinterpret_cast<uintptr_t>(s.handle)); });
    py::class_<XYZ>(c, "xyz", py::dynamic_attr())
        .def(py::init<F64, F64, F64>(), "x"_a, "y"_a, "z"_a)
        .def_readwrite("x", &XYZ::X).def_readwrite("y", &XYZ::Y).def_readwrite("z", &XYZ::Z);
    py::class_<UVW>(c, "uvw", py::dynamic_attr())
        .def(py::init<F64, F64, F64>(), "u"_a, "v"_a, "w"_a)
        .def_readwrite("u", &UVW::U).def_readwrite("v", &UVW::V).def_readwrite("w", &UVW::W);
    py::class_<LLA>(c, "lla", py::dynamic_attr())
        .def(py::init<F64, F64, F64>(), "lon"_a, "lat"_a, "alt"_a)
        .def_readwrite("lon", &LLA::LON).def_readwrite("lat", &LLA::LAT).def_readwrite("alt", &LLA::ALT);
    py::class_<RA_DEC>(c, "ra_dec", py::dynamic_attr())
        .def(py::init<F64, F64>(), "ra"_a, "dec"_a)
        .def_readwrite("ra", &RA_DEC::RA).def_readwrite("dec", &RA_DEC::DEC);
    c.def("set_device", [](int d) { return SetCudaDevice(d); }, "device_id"_a);

    // --------------------------------------------------------- _hidden_impl
    auto h = m.def_submodule("_hidden_impl");
    py::class_<Module, std::shared_ptr<Module>>(h, "module")
        .def("name", &Module::name)
        .def("process", [](Module& mod, U64 step) { return mod.process(step); }, "step"_a = 0);
    py::class_<Bundle, std::shared_ptr<Bundle>>(h, "bundle")
        .def("modules", &Bundle::getModules);

    // ------------------------------------------------------------ _mem_impl
    auto mem = m.def_submodule("_mem_impl");
    py::class_<ArrayShape>(mem, "array_shape")
        .def(py::init<const ArrayShape::Type&>(), "shape"_a)
        .def_property_readonly("number_of_aspects", &ArrayShape::numberOfAspects)
        .def_property_readonly("number_of_frequency_channels", &ArrayShape::numberOfFrequencyChannels)
        .def_property_readonly("number_of_time_samples", &ArrayShape::numberOfTimeSamples)
        .def_property_readonly("number_of_polarizations", &ArrayShape::numberOfPolarizations)
        .def("__getitem__", [](ArrayShape& s, U64 i) { return shape_at(s, i); })
        .def("__len__", [](ArrayShape&) { return 4; })
        .def("__repr__", [](ArrayShape& s) { return "ArrayShape(shape=" + shape_str(s) + ")"; });
    py::class_<PhasorShape>(mem, "phasor_shape")
        .def(py::init<const PhasorShape::Type&>(), "shape"_a)
        .def_property_readonly("number_of_beams", &PhasorShape::numberOfBeams)
        .def_property_readonly("number_of_antennas", &PhasorShape::numberOfAntennas)
        .def_property_readonly("number_of_frequency_channels", &PhasorShape::numberOfFrequencyChannels)
        .def_property_readonly("number_of_time_samples", &PhasorShape::numberOfTimeSamples)
        .def_property_readonly("number_of_polarizations", &PhasorShape::numberOfPolarizations)
        .def("__getitem__", [](PhasorShape& s, U64 i) { return shape_at(s, i); })
        .def("__len__", [](PhasorShape&) { return 5; })
        .def("__repr__", [](PhasorShape& s) { return "PhasorShape(shape=" + shape_str(s) + ")"; });
    py::class_<DelayShape>(mem, "delay_shape")
        .def(py::init<const DelayShape::Type&>(), "shape"_a)
        .def_property_readonly("number_of_beams", &DelayShape::numberOfBeams)
        .def_property_readonly("number_of_antennas", &DelayShape::numberOfAntennas)
        .def("__getitem__", [](DelayShape& s, U64 i) { return shape_at(s, i); })
        .def("__len__", [](DelayShape&) { return 2; })
        .def("__repr__", [](DelayShape& s) { return "DelayShape(shape=" + shape_str(s) + ")"; });
    py::class_<VectorShape>(mem, "vector_shape")
        .def(py::init<const VectorShape::Type&>(), "shape"_a)
        .def("__getitem__", [](VectorShape& s, U64 i) { return shape_at(s, i); })
        .def("__len__", [](VectorShape&) { return 1; })
        .def("__repr__", [](VectorShape& s) { return "VectorShape(shape=" + shape_str(s) + ")"; });
    bind_device<Device::CPU>(mem, "cpu");
    bind_device<Device::CUDA>(mem, "cuda");

    // ----------------------------------------------------------- _copy_impl
    // Byte copies between any two tensors; the Python layer resolves tensor
    // types, duet slots and shape checks (reference python/blade/ext_copy.cc).
    auto cp = m.def_submodule("_copy_impl");
    cp.def("copy_bytes", [](uintptr_t dst, uintptr_t src, U64 nbytes, uintptr_t stream, bool sync) {
        if (nbytes == 0 || dst == src) return Result::SUCCESS;
        hipStream_t s = reinterpret_cast<hipStream_t>(stream);
        // Host to host: finished on return, like Copy (include/blade/memory/copy.hh).
        if (is_host_memory(reinterpret_cast<const void*>(dst)) && is_host_memory(reinterpret_cast<const void*>(src))) {
            if (hipStreamSynchronize(s) != hipSuccess) return Result::CUDA_ERROR;
            std::memcpy(reinterpret_cast<void*>(dst), reinterpret_cast<const void*>(src), nbytes);
            return Result::SUCCESS;
        }
        hipError_t e = hipMemcpyAsync(reinterpret_cast<void*>(dst), reinterpret_cast<const void*>(src), nbytes,
                                      hipMemcpyDefault, s);
        if (e == hipSuccess && sync) e = hipStreamSynchronize(s);
        return e == hipSuccess ? Result::SUCCESS : Result::CUDA_ERROR;
    }, "dst"_a, "src"_a, "nbytes"_a, "stream"_a = 0, "sync"_a = true);

    // ------------------------------------------------ _pipeline / _runner
    auto pm = m.def_submodule("_pipeline_impl");
    py::class_<Pipeline> pcls(pm, "pipeline", py::dynamic_attr());
    pcls.def(py::init<>());
    bind_pipeline_methods(pcls);

    auto rm = m.def_submodule("_runner_impl");
    py::class_<Runner> rcls(rm, "runner", py::dynamic_attr());
    rcls.def(py::init<>())
        .def("enqueue", &Runner::enqueue, "input_callback"_a, "output_callback"_a, "input_id"_a, "output_id"_a)
        .def("dequeue", &Runner::dequeue, "callback"_a)
        .def("head", &Runner::head)
        .def("queued", &Runner::queued);
    bind_pipeline_methods(rcls);

    // -------------------------------------------------------------- modules
    auto bf = m.def_submodule("_beamformer_impl");
    bind_beamformer<CF32, CF32>(bf, "in_cf32", "out_cf32");
    bind_beamformer<CI8, CF32>(bf, "in_ci8", "out_cf32");

    auto bb = m.def_submodule("_batchbeamformer_impl");
    bind_batched_beamformer(bb);

    auto ch = m.def_submodule("_channelizer_impl");
    bind_channelizer<CF32, CF32>(ch, "in_cf32", "out_cf32");

    auto po = m.def_submodule("_polarizer_impl");
    bind_polarizer<CF32, CF32>(po, "in_cf32", "out_cf32");
    bind_polarizer<CF16, CF16>(po, "in_cf16", "out_cf16");

    auto de = m.def_submodule("_detector_impl");
    bind_detector<CF32, F32>(de, "in_cf32", "out_f32");

    // The reference's 22 pairs (python/blade/ext_cast.cc:43-75); the pairs
    // between real and complex types throw at construction, as there.
    auto ca = m.def_submodule("_cast_impl");
    bind_cast<I8, I8>(ca, "in_i8", "out_i8");
    bind_cast<I8, F32>(ca, "in_i8", "out_f32");
    bind_cast<F16, F32>(ca, "in_f16", "out_f32");
    bind_cast<F16, CF32>(ca, "in_f16", "out_cf32");
    bind_cast<F16, CF16>(ca, "in_f16", "out_cf16");
    bind_cast<F16, F16>(ca, "in_f16", "out_f16");
    bind_cast<F32, I32>(ca, "in_f32", "out_i32");
    bind_cast<F32, F16>(ca, "in_f32", "out_f16");
    bind_cast<F32, CF32>(ca, "in_f32", "out_cf32");
    bind_cast<F32, CF16>(ca, "in_f32", "out_cf16");
    bind_cast<F32, F32>(ca, "in_f32", "out_f32");
    bind_cast<CI8, CF32>(ca, "in_ci8", "out_cf32");
    bind_cast<CI8, CF16>(ca, "in_ci8", "out_cf16");
    bind_cast<CI8, CI8>(ca, "in_ci8", "out_ci8");
    bind_cast<CF32, F16>(ca, "in_cf32", "out_f16");
    bind_cast<CF32, F32>(ca, "in_cf32", "out_f32");
    bind_cast<CF32, CF16>(ca, "in_cf32", "out_cf16");
    bind_cast<CF32, CF32>(ca, "in_cf32", "out_cf32");
    bind_cast<CF16, F16>(ca, "in_cf16", "out_f16");
    bind_cast<CF16, F32>(ca, "in_cf16", "out_f32");
    bind_cast<CF16, CF32>(ca, "in_cf16", "out_cf32");
    bind_cast<CF16, CF16>(ca, "in_cf16", "out_cf16");

    auto ga = m.def_submodule("_gather_impl");
    bind_gather<CI8, CI8>(ga, "in_ci8", "out_ci8");
    bind_gather<CF16, CF16>(ga, "in_cf16", "out_cf16");
    bind_gather<CF32, CF32>(ga, "in_cf32", "out_cf32");
    bind_gather<F32, F32>(ga, "in_f32", "out_f32");
    // The reference registers the F32 gather as in_f64 / out_f64
    // (python/blade/ext_gather.cc:50); both paths resolve to it here.
    ga.def_submodule("in_f64").attr("out_f64") = ga.attr("in_f32").attr("out_f32");

    auto du = m.def_submodule("_duplicate_impl");
    bind_duplicate<CI8, CI8>(du, "in_ci8", "out_ci8");
    bind_duplicate<CF16, CF16>(du, "in_cf16", "out_cf16");
    bind_duplicate<CF32, CF32>(du, "in_cf32", "out_cf32");
    bind_duplicate<F32, F32>(du, "in_f32", "out_f32");

    auto gu = m.def_submodule("_guppi_impl");
    auto gr = gu.def_submodule("taint_reader");
    bind_guppi_reader<CI8>(gr, "out_ci8", "type_ci8");
    auto gw = gu.def_submodule("taint_writer");
    bind_guppi_writer<CF16>(gw, "in_cf16", "type_cf16");
    bind_guppi_writer<CF32>(gw, "in_cf32", "type_cf32");

    auto pe = m.def_submodule("_permutation_impl");
    bind_permutation<CI8, CI8>(pe, "in_ci8", "out_ci8");
    bind_permutation<CF16, CF16>(pe, "in_cf16", "out_cf16");
    bind_permutation<CF32, CF32>(pe, "in_cf32", "out_cf32");

    auto ph = m.def_submodule("_phasor_impl");
    bind_phasor<CF32>(ph, "out_cf32");
    bind_phasor<CF64>(ph, "out_cf64");

    auto mb = m.def_submodule("_modeb_impl");
    bind_modeb<CI8, CF32>(mb, "in_ci8", "out_cf32");
    bind_modeb<CI8, F32>(mb, "in_ci8", "out_f32");
    bind_modeb<CF32, CF32>(mb, "in_cf32", "out_cf32");
    bind_modeb<CF32, F32>(mb, "in_cf32", "out_f32");

    auto mh = m.def_submodule("_modeh_impl");
    bind_modeh<CF32, F32>(mh, "in_cf32", "out_f32");
    bind_modeh<CI8, F32>(mh, "in_ci8", "out_f32");
}
