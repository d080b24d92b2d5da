// Pipeline, Runner and Profiler of the Blade C++ API on HIP streams
// (reference src/pipeline.cc:7-141, src/runner.cc:17-68, src/memory/profiler.cc).
#include "blade/base.hh"

#include <rocprofiler-sdk-roctx/roctx.h>

#include <cstdlib>
#include <thread>

namespace Blade {

namespace {
// BLADE_ROCTX=1: every module's process() call becomes a ROCTX range named
// after the module ("rocprofv3 --marker-trace" shows them beside the kernels).
bool markersEnabled() {
    static const bool on = [] {
        const char* v = std::getenv("BLADE_ROCTX");
        return v && std::atoi(v) != 0;
    }();
    return on;
}
}  // namespace

#undef BL_LOG_DOMAIN
#define BL_LOG_DOMAIN "PIPELINE"

Profiler& Profiler::Get() {
    static Profiler instance;
    return instance;
}

Pipeline::Pipeline()
    : _commited(false), _graphs(false), _computeStepCount(0), _computeStepsPerCycle(1),
      _computeLifetimeCycles(0), _cycleDone(nullptr), _cycleRecorded(false) {
    _streams.resize(2);
    _graphExec.assign(2, nullptr);
    for (auto& s : _streams) {
        // Default priority; non-blocking with respect to the null stream.
        BL_HIP_CHECK_THROW(hipStreamCreateWithFlags(&s.handle, hipStreamNonBlocking));
    }
    BL_HIP_CHECK_THROW(hipEventCreateWithFlags(&_cycleDone, hipEventDisableTiming));
}

Pipeline::~Pipeline() {
    for (U64 i = 0; i < _streams.size(); i++) {
        (void)synchronize(i);
    }
    for (auto& g : _graphExec) {
        if (g) (void)hipGraphExecDestroy(g);
    }
    _modules.clear();
    if (_cycleDone) (void)hipEventDestroy(_cycleDone);
    for (auto& s : _streams) (void)hipStreamDestroy(s.handle);
}

Result Pipeline::markCycleDone(const U64& index) {
    BL_HIP_CHECK(hipEventRecord(_cycleDone, _streams[index].handle));
    _cycleRecorded = true;
    return Result::SUCCESS;
}

void Pipeline::addModule(const std::shared_ptr<Module>& module) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        BL_FATAL("Error while creating module '{}' in position '{}': '{}'", module->name(),
                 _modules.size(), hipGetErrorString(e));
        throw Result::CUDA_ERROR;
    }
    if ((module->getTaint() & Taint::CHRONOUS) == Taint::CHRONOUS) {
        const U64 ratio = module->getComputeRatio();
        if (ratio > 1) {
            _computeStepsPerCycle *= ratio;
            _computeStepRatios.push_back(ratio);
        }
    }
    _modules.push_back(module);
}

Result Pipeline::synchronize(const U64& index) {
    BL_HIP_CHECK(hipStreamSynchronize(_streams[index].handle));
    return Result::SUCCESS;
}

bool Pipeline::isSynchronized(const U64& index) {
    return hipStreamQuery(_streams[index].handle) == hipSuccess;
}

namespace {

// Beamformer -> [Channelizer of rate 1, 2 or 4] -> [Polarizer], each reading
// and writing the beamformer's output buffer in place and adjacent in the
// module order, become one launch of the beamformer with a fused epilogue.
// The buffer ends up as after the three modules ran one by one (the
// channelizer and polarizer alias it like the reference's Link,
// src/modules/channelizer/base.cc:45, polarizer/base.cc:54), so the fusion
// is invisible except in the launch count.
template <typename IT>
bool fuseBeamformerChain(const std::vector<std::shared_ptr<Module>>& mods, U64 i) {
    auto* bf = dynamic_cast<Modules::Beamformer::Generic<IT, CF32>*>(mods[i].get());
    if (!bf) return false;
    const void* buf = bf->getOutputBuffer().data();
    U64 j = i + 1;
    U64 rate = 1;
    Modules::Channelizer<CF32, CF32>* ch = nullptr;
    Modules::Polarizer<CF32, CF32>* po = nullptr;
    if (j < mods.size()) {
        ch = dynamic_cast<Modules::Channelizer<CF32, CF32>*>(mods[j].get());
        if (ch && ch->getInputBuffer().data() == buf && ch->getOutputBuffer().data() == buf &&
            ch->getConfig().rate <= 4 && ch->getConfig().rate != 3) {
            rate = ch->getConfig().rate;
            j += 1;
        } else {
            ch = nullptr;
        }
    }
    if (j < mods.size()) {
        po = dynamic_cast<Modules::Polarizer<CF32, CF32>*>(mods[j].get());
        if (!(po && po->getInputBuffer().data() == buf && po->getOutputBuffer().data() == buf &&
              po->getConfig().inputPolarization != po->getConfig().outputPolarization)) {
            po = nullptr;
        }
    }
    if (rate == 1 && !po) return false;       // nothing to fuse
    if (bf->fuseEpilogue(rate, po != nullptr) != Result::SUCCESS) return false;
    if (ch) ch->fuseIntoProducer();
    if (po) po->fuseIntoProducer();
    BL_INFO("Fused Beamformer{}{} into one launch.", ch ? " -> Channelizer" : "",
            po ? " -> Polarizer" : "");
    return true;
}

}  // namespace

Result Pipeline::commit() {
    BL_DEBUG("Commiting pipeline with {} modules and {} compute steps per cycle.", _modules.size(),
             _computeStepsPerCycle);
    for (U64 i = 0; i < _modules.size(); i++) {
        if (!fuseBeamformerChain<CI8>(_modules, i)) (void)fuseBeamformerChain<CF32>(_modules, i);
    }
    _computeStepRatios.push_back(1);
    return Result::SUCCESS;
}

// A captured graph replays the device work of one call: every module must
// run every cycle (no CHRONOUS step ratios) and do no host work in process().
Result Pipeline::checkCapturable() const {
    if (_computeStepsPerCycle != 1) {
        BL_FATAL("Graph capture needs a pipeline without CHRONOUS modules.");
        return Result::ERROR;
    }
    for (const auto& module : _modules) {
        if (!module->graphCapturable()) {
            BL_FATAL("Graph capture: module '{}' does host work in process().", module->name());
            return Result::ERROR;
        }
    }
    return Result::SUCCESS;
}

Result Pipeline::enableGraphCapture(bool enable) {
    if (enable) BL_CHECK(checkCapturable());
    _graphs = enable;
    return Result::SUCCESS;
}

// Run every module once on stream `index`, with the step counters of the
// reference's CHRONOUS scheduling: module k sees the step count of its
// cycle, and modules after a CHRONOUS module run only when it completes.
Result Pipeline::runModules(const U64& index) {
    U64 localRatioIndex = 0;
    U64 localStepOffset = 1;
    for (auto& module : _modules) {
        const U64 ratio = _computeStepRatios[localRatioIndex];
        const U64 localStepCount = (_computeStepCount / localStepOffset) % ratio;
        BL_TRACE("[M: '{}', R: {}]: Step Count: {}, Step Offset: {}, Step Ratio: {}", module->name(),
                 module->getComputeRatio(), localStepCount, localStepOffset, ratio);
        if (markersEnabled()) roctxRangePushA(module->name().c_str());
        const Result result = module->process(localStepCount, _streams[index]);
        if (markersEnabled()) roctxRangePop();
        if (result == Result::PIPELINE_EXHAUSTED) {
            BL_INFO("Module finished pipeline execution at {} lifetime compute cycles.",
                    _computeLifetimeCycles);
            return Result::PIPELINE_EXHAUSTED;
        }
        if (result != Result::SUCCESS) return result;
        if (module->getComputeRatio() > 1) {
            if (localStepCount == ratio - 1) {
                localStepOffset *= ratio;
                localRatioIndex += 1;
            } else {
                break;
            }
        }
    }
    return Result::SUCCESS;
}

Result Pipeline::compute(const U64& index) {
    if (!_commited) {
        BL_CHECK(commit());
        _commited = true;
    }
    // The previous cycle (on either stream) must be over before this one
    // writes the modules' shared outputs.
    if (_cycleRecorded) BL_HIP_CHECK(hipStreamWaitEvent(_streams[index].handle, _cycleDone, 0));
    if (_graphs) {
        hipStream_t s = _streams[index].handle;
        if (!_graphExec[index]) {
            // Modules connected after enableGraphCapture() are checked here.
            BL_CHECK(checkCapturable());
            hipGraph_t graph = nullptr;
            BL_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            const Result r = runModules(index);
            const hipError_t e = hipStreamEndCapture(s, &graph);
            if (r != Result::SUCCESS) {
                if (graph) (void)hipGraphDestroy(graph);
                return r;
            }
            BL_HIP_CHECK(e);
            BL_HIP_CHECK(hipGraphInstantiate(&_graphExec[index], graph, nullptr, nullptr, 0));
            (void)hipGraphDestroy(graph);
        }
        BL_HIP_CHECK(hipGraphLaunch(_graphExec[index], s));
    } else {
        const Result r = runModules(index);
        if (r != Result::SUCCESS) return r;
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        BL_FATAL("HIP compute error: {}", hipGetErrorString(e));
        return Result::CUDA_ERROR;
    }
    BL_CHECK(markCycleDone(index));
    _computeStepCount += 1;
    if (_computeStepCount == _computeStepsPerCycle) {
        _computeStepCount = 0;
        _computeLifetimeCycles += 1;
    }
    return Result::SUCCESS;
}

#undef BL_LOG_DOMAIN
#define BL_LOG_DOMAIN "RUNNER"

Runner::Runner() : headIndex(0) {}

Runner::~Runner() {
    if (!queue.empty()) BL_WARN("Runner destroyed with {} stream left inside queue.", queue.size());
}

Result Runner::enqueue(const std::function<Result()>& inputCallback,
                       const std::function<Result()>& outputCallback, U64 inputId, U64 outputId) {
    if (queue.size() == numberOfStreams()) {
        std::this_thread::yield();
        return Result::RUNNER_QUEUE_FULL;
    }
    const bool out = willOutput();
    BL_TRACE("[E] Index: {} | Input ID: {} | Output ID: {} | Will Output: {}", headIndex, inputId,
             outputId, out ? "Y" : "N");
    BL_CHECK(inputCallback());
    BL_CHECK(compute(headIndex));
    if (out) {
        BL_CHECK(outputCallback());
        // The next cycle's compute also waits for these output copies.
        BL_CHECK(markCycleDone(headIndex));
    }
    queue.push({headIndex, inputId, outputId, out});
    headIndex = (headIndex + 1) % numberOfStreams();
    return Result::SUCCESS;
}

Result Runner::dequeue(const DequeueCallback& callback) {
    if (queue.empty()) {
        std::this_thread::yield();
        return Result::RUNNER_QUEUE_EMPTY;
    }
    const auto [frontIndex, inputId, outputId, didOutput] = queue.front();
    if (queue.size() < numberOfStreams()) {
        if (!isSynchronized(frontIndex)) return Result::RUNNER_QUEUE_NONE_AVAILABLE;
    } else {
        BL_CHECK(synchronize(frontIndex));
    }
    queue.pop();
    BL_TRACE("[D] Index: {} | Input ID: {} | Output ID: {} | Did Output: {}", frontIndex, inputId,
             outputId, didOutput ? "Y" : "N");
    return callback(inputId, outputId, didOutput);
}

}  // namespace Blade
