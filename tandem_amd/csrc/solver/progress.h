// Progress lines of long solves on stderr, at most one every TDG_PROGRESS_SECONDS seconds (the
// environment variable; unset or 0: none).  Printed from the solver's own host checks.
#pragma once

#include <chrono>
#include <cstdio>
#include <cstdlib>

namespace tdg {

class Progress {
public:
    explicit Progress(char const* what) : what_(what) {
        char const* e = std::getenv("TDG_PROGRESS_SECONDS");
        period_ = e ? std::atof(e) : 0.0;
        last_ = std::chrono::steady_clock::now();
    }
    // it: iterations so far; active: systems still iterating; res: a residual figure to report
    void tick(int it, int active, double res) {
        if (!(period_ > 0.0)) return;
        const auto now = std::chrono::steady_clock::now();
        if (std::chrono::duration<double>(now - last_).count() < period_) return;
        last_ = now;
        std::fprintf(stderr, "[tdg] %s: iteration %d, %d active, residual %.3e\n", what_, it, active, res);
        std::fflush(stderr);
    }

private:
    char const* what_;
    double period_ = 0.0;
    std::chrono::steady_clock::time_point last_;
};

} // namespace tdg
