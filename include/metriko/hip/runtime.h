// metriko (MI355X build) -- process-wide HIP context used by the drop-in headers.
//
// The reference's solver seams (solveSmallestEig, iterative_rounding, ...) take no
// context argument, so the headers share one lazily created mk_ctx on device 0
// (METRIKO_DEVICE overrides).  Options default to FP32 inner solves with FP64
// refinement to a 1e-10 relative residual; set_options() changes them.
#ifndef METRIKO_HIP_RUNTIME_H
#define METRIKO_HIP_RUNTIME_H
#include <cstdlib>
#include <stdexcept>
#include <string>
#include "metriko_hip.h"

namespace metriko { namespace hip {
    struct Runtime {
        mk_ctx* ctx = nullptr;
        mk_solve_opts field_opts{};
        mk_solve_opts igm_opts{};
        mk_stats last_field{};
        mk_igm_stats last_igm{};
        bool external = false;
        Runtime() {
            mk_default_opts(&field_opts);
            mk_default_opts(&igm_opts);
        }
        ~Runtime() { if (ctx && !external) mk_finalize(ctx); }
        mk_ctx* get() {
            if (!ctx) {
                const char* d = std::getenv("METRIKO_DEVICE");
                mk_status s = mk_init(&ctx, d ? std::atoi(d) : 0);
                if (s != MK_OK) throw std::runtime_error(std::string("metriko: mk_init failed: ") + mk_status_str(s));
            }
            return ctx;
        }
        // use a context owned by the caller (the C driver entry point does this)
        void attach(mk_ctx* c) { if (ctx && !external) mk_finalize(ctx); ctx = c; external = true; }
    };
    inline Runtime& runtime() { static thread_local Runtime r; return r; }
    inline void check(mk_status s, const char* what) {
        if (s != MK_OK) {
            char buf[512] = {0};
            if (runtime().ctx) mk_last_error(runtime().ctx, buf, sizeof(buf));
            throw std::runtime_error(std::string("metriko: ") + what + ": " + mk_status_str(s) + " " + buf);
        }
    }
}}

#endif
