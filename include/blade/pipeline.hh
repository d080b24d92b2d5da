// Pipeline: an ordered list of modules executed on one of two HIP streams,
// with CHRONOUS step-ratio scheduling (reference include/blade/pipeline.hh:15-89,
// src/pipeline.cc:7-141).  Optionally each stream's module sequence is
// captured once into a hipGraph and replayed (pipelines without CHRONOUS
// modules and without host-side work in process()).
//
// Cross-stream order.  Every module owns one output buffer, shared by both
// streams, and CHRONOUS modules (Gather, a stepped Detector) accumulate into
// theirs across cycles.  So a cycle's compute waits on the device for the
// previous cycle to end: its compute, and under a Runner its output copies.
// The input copies of the next cycle (into the other slot of a Duet) are
// not ordered and overlap the previous cycle's compute.
#ifndef BLADE_PIPELINE_HH
#define BLADE_PIPELINE_HH

#include <memory>
#include <type_traits>
#include <vector>

#include "blade/bundle.hh"
#include "blade/macros.hh"
#include "blade/module.hh"

namespace Blade {

class BLADE_API Pipeline {
 public:
    Pipeline();
    virtual ~Pipeline();

    bool computeComplete() const { return (_computeStepCount + 1) == _computeStepsPerCycle; }
    const U64& computeCurrentStepCount() const { return _computeStepCount; }
    const U64& computeStepsPerCycle() const { return _computeStepsPerCycle; }
    const U64& computeLifetimeCycles() const { return _computeLifetimeCycles; }
    const bool& commited() const { return _commited; }
    bool willOutput() const { return computeStepsPerCycle() == (computeCurrentStepCount() + 1); }

    template <typename Block>
    void connect(std::shared_ptr<Block>& module, const typename Block::Config& config,
                 const typename Block::Input& input) {
        if (_commited) {
            BL_FATAL("Can't connect new module after Pipeline is commited.");
            throw Result::ERROR;
        }
        module = std::make_shared<Block>(config, input, _streams[0]);
        if constexpr (std::is_base_of<Bundle, Block>::value) {
            for (auto& mod : module->getModules()) addModule(mod);
        } else {
            addModule(module);
        }
    }

    void addModule(const std::shared_ptr<Module>& module);

    Result compute(const U64& index);
    Result synchronize(const U64& index);
    bool isSynchronized(const U64& index);

    // Capture each stream's module sequence into a hipGraph on its first
    // compute() and replay it afterwards.
    Result enableGraphCapture(bool enable = true);
    bool graphCaptureEnabled() const { return _graphs; }

    const Stream& stream(const U64& index = 0) const { return _streams[index]; }
    U64 numberOfStreams() const { return _streams.size(); }
    const std::vector<std::shared_ptr<Module>>& modules() const { return _modules; }

 private:
    bool _commited;
    bool _graphs;
    std::vector<Stream> _streams;
    std::vector<std::shared_ptr<Module>> _modules;
    std::vector<hipGraphExec_t> _graphExec;

    U64 _computeStepCount;
    U64 _computeStepsPerCycle;
    U64 _computeLifetimeCycles;
    std::vector<U64> _computeStepRatios;

    // Recorded at the end of the last cycle (compute, or the Runner's output
    // callback); the next compute on any stream waits on it.
    hipEvent_t _cycleDone;
    bool _cycleRecorded;

    Result commit();
    Result runModules(const U64& index);
    Result checkCapturable() const;

 protected:
    // Mark the work issued so far on stream `index` as the end of the cycle.
    Result markCycleDone(const U64& index);
};

}  // namespace Blade

#endif
