// Replays a fixed sequence of launches: captured once into a HIP graph on a non-default stream
// (solver iterations between two host checks are the same launches with the same arguments), or
// launched directly when graphs are off or the stream is the default stream, which cannot be
// captured.  The sequence must not allocate or memset.
#pragma once

#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

namespace tdg {

class Replay {
public:
    Replay(hipStream_t s, bool enable) : s_(s), on_(enable && s != nullptr) {}
    Replay(Replay const&) = delete;
    Replay& operator=(Replay const&) = delete;
    ~Replay() {
        if (exec_) (void)hipGraphExecDestroy(exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
    }
    template <class F> void run(F&& launches) {
        if (!on_) {
            launches();
            return;
        }
        if (!exec_) {
            check(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
            try {
                launches();
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(s_, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            check(hipStreamEndCapture(s_, &graph_));
            check(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
        }
        check(hipGraphLaunch(exec_, s_));
    }

private:
    static void check(hipError_t e) {
        if (e != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e));
    }
    hipStream_t s_;
    bool on_;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
};

} // namespace tdg
