// Internal interface of the never-stored cost (fused.hip) for the solvers that
// drive it: the fixed-iteration schedule (fused.hip) and the duality-gap
// schedule (sinkhorn.hip).  Not part of the C ABI.
#pragma once
#include <cstddef>
#include <cstdint>

#include <hip/hip_runtime.h>

namespace wotk {

struct FzSelect;

// C = sqeuclid(X, Y) / median(sqeuclid(X, Y)) regenerated per tile from the
// scaled, centred coordinates' split-BF16 pieces; buffers carved from a
// caller workspace by fused_problem_setup.
struct FusedProblem {
    const double *X = nullptr, *Y = nullptr;
    int I = 0, J = 0, d = 0, kp = 0, terms = 6;
    double *mean = nullptr, *xs = nullptr, *ys = nullptr, *nx = nullptr, *ny = nullptr, *zeros = nullptr;
    __bf16 *xp[3] = {nullptr, nullptr, nullptr}, *yp[3] = {nullptr, nullptr, nullptr};
    int *bad = nullptr, *nbad = nullptr;    // repair list of the sweeps
    int *bad2 = nullptr, *nbad2 = nullptr;  // and of the sum passes (which may run on another stream)
    double *part = nullptr, *part2 = nullptr;
    FzSelect *sel = nullptr;
    unsigned long long *hist = nullptr;
    const double *median = nullptr;  // device: the exact median of the FP64 distances
};

// workspace bytes of the problem buffers (taken first from an arena)
size_t fused_problem_bytes(int64_t I, int64_t J, int64_t d);

// carve the buffers at the start of `workspace`, compute the exact median and
// the split coordinates; returns the bytes used in *used
int fused_problem_setup(const double *X, const double *Y, int64_t I, int64_t J, int64_t d, int terms,
                        void *workspace, size_t workspace_bytes, FusedProblem *fp, size_t *used, hipStream_t s);

// one (a, b) iteration at eps: f from g (source rows, log weights logp), then
// g from f (target rows, log q at logq_dev[0])
int fused_problem_iterate(const FusedProblem &fp, double *f, double *g, const double *logp,
                          const double *logq_dev, double lambda1, double lambda2, double eps, hipStream_t s);

// sums of exp((f_i + g_j - C_ij)/eps) / div over the targets of every source
// (over_targets false: out has I entries) or over the sources of every target
// (true: J entries).  exact: FP64 exponentials and sums (the gap pass), else
// the sweeps' arithmetic.  Uses the second repair list: it may run on another
// stream than fused_problem_iterate.
int fused_problem_sums(const FusedProblem &fp, bool over_targets, const double *f, const double *g, double eps,
                       double div, bool exact, double *out, hipStream_t s);

}  // namespace wotk
