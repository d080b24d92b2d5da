// Pipeline, Runner and Profiler of the Blade C++ API on HIP streams
// (reference src/pipeline.cc:7-141, src/runner.cc:17-68, src/memory/profiler.cc).
#include "blade/base.hh"

#include <thread>

namespace Blade {

#undef BL_LOG_DOMAIN
#define BL_LOG_DOMAIN "PIPELINE"

Profiler& Profiler::Get() {
    static Profiler instance;
    return instance;
}

Pipeline::Pipeline()
    : _commited(false), _graphs(false), _computeStepCount(0), _computeStepsPerCycle(1),
      _computeLifetimeCycles(0) {
    _streams.resize(2);
    _graphExec.assign(2, nullptr);
    for (auto& s : _streams) {
        // Default priority; non-blocking with respect to the null stream.
        BL_HIP_CHECK_THROW(hipStreamCreateWithFlags(&s.handle, hipStreamNonBlocking));
    }
}

Pipeline::~Pipeline() {
    for (U64 i = 0; i < _streams.size(); i++) {
        (void)synchronize(i);
    }
    for (auto& g : _graphExec) {
        if (g) (void)hipGraphExecDestroy(g);
    }
    _modules.clear();
    for (auto& s : _streams) (void)hipStreamDestroy(s.handle);
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

Result Pipeline::commit() {
    BL_DEBUG("Commiting pipeline with {} modules and {} compute steps per cycle.", _modules.size(),
             _computeStepsPerCycle);
    _computeStepRatios.push_back(1);
    return Result::SUCCESS;
}

Result Pipeline::enableGraphCapture(bool enable) {
    if (enable && _computeStepsPerCycle != 1) {
        BL_FATAL("Graph capture needs a pipeline without CHRONOUS modules.");
        return Result::ERROR;
    }
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
        const Result result = module->process(localStepCount, _streams[index]);
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
    if (_graphs) {
        hipStream_t s = _streams[index].handle;
        if (!_graphExec[index]) {
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
    if (out) BL_CHECK(outputCallback());
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
