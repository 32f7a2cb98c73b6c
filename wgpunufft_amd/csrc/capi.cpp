// C ABI of the engine: the device-pointer API (include/mnufft.h) and the
// host-array API compatible with the reference's C interface
// (include/wgpu_nufft.h; reference capi/src/lib.rs:326-934,
// capi/src/plan.rs:156-893, capi/src/error.rs:27-91).
//
// No C++ exception or panic crosses the boundary: every entry point catches
// and maps failures to the documented codes, with a thread-local message.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mnufft.h"
#include "../../include/wgpu_nufft.h"
#include "modes.hpp"
#include "plan.hpp"

using mnufft::Plan;
using mnufft::Precision;
using mnufft::set_error;

namespace {

constexpr double kMaxCoordinate = 3.0 * M_PI;


bool gpu_available(std::string *why)
{
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        if (why) *why = "no HIP device is available";
        return false;
    }
    return true;
}

template <typename F>
int guarded(F &&f)
{
    try {
        return f();
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return MNUFFT_ERROR_INTERNAL;
    } catch (const std::exception &e) {
        set_error(std::string("internal error: ") + e.what());
        return MNUFFT_ERROR_INTERNAL;
    } catch (...) {
        set_error("internal error");
        return MNUFFT_ERROR_INTERNAL;
    }
}

int make_plan(int type, int dim, const int64_t *n_modes, int isign, int64_t batch, double eps,
              int precision, int mode_order, double sigma, std::unique_ptr<Plan> *out)
{
    if (type < 1 || type > 2) {
        set_error("plan type must be 1 or 2 here (type 3 has its own constructor)");
        return MNUFFT_ERROR_INVALID_ARGUMENT;
    }
    if (dim < 1 || dim > mnufft::kMaxDims) {
        set_error("GPU plans take 1 to 8 dimensions, got " + std::to_string(dim));
        return MNUFFT_ERROR_PLAN;
    }
    if (!n_modes) {
        set_error("n_modes is null");
        return MNUFFT_ERROR_INVALID_ARGUMENT;
    }
    if (batch < 1) {
        set_error("batch must be at least 1");
        return MNUFFT_ERROR_PLAN;
    }
    if (mode_order != 0 && mode_order != 1) {
        set_error("mode order must be 0 (centered) or 1 (FFT)");
        return MNUFFT_ERROR_INVALID_ARGUMENT;
    }
    std::string why;
    if (!gpu_available(&why)) {
        set_error(why);
        return MNUFFT_ERROR_GPU_UNAVAILABLE;
    }
    auto p = std::make_unique<Plan>();
    p->type = type;
    p->dim = dim;
    for (int a = 0; a < dim; ++a) {
        if (n_modes[a] < 1) {
            set_error("mode count along axis " + std::to_string(a) + " must be positive");
            return MNUFFT_ERROR_PLAN;
        }
        p->n_modes[a] = n_modes[a];
    }
    p->isign = isign >= 0 ? 1 : -1;
    p->batch = batch;
    p->eps = eps;
    p->sigma = sigma == 0.0 ? 2.0 : sigma;
    p->precision = precision == MNUFFT_PRECISION_F32 ? Precision::F32
                   : precision == MNUFFT_PRECISION_DF64 ? Precision::Df64
                                                        : Precision::F64;
    p->mode_order = mode_order == 1 ? mnufft::ModeOrder::Fft : mnufft::ModeOrder::Centered;
    if (!p->init_common()) return MNUFFT_ERROR_PLAN;
    *out = std::move(p);
    return MNUFFT_SUCCESS;
}

}  // namespace

// ======================================================================
// Device-pointer API.
// ======================================================================
struct mnufft_plan_s {
    std::unique_ptr<Plan> plan;
};

extern "C" {

const char *mnufft_last_error(void) { return mnufft::last_error(); }

int mnufft_makeplan(int32_t type, int32_t dim, const int64_t *n_modes, int32_t isign,
                    int64_t batch, double eps, int32_t precision, int32_t mode_order, double sigma,
                    mnufft_plan *plan)
{
    return guarded([&] {
        if (!plan) {
            set_error("plan output pointer is null");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        *plan = nullptr;
        std::unique_ptr<Plan> p;
        int rc = make_plan(type, dim, n_modes, isign, batch, eps, precision, mode_order, sigma, &p);
        if (rc) return rc;
        *plan = new mnufft_plan_s{std::move(p)};
        return MNUFFT_SUCCESS;
    });
}

int mnufft_makeplan_type3(int32_t dim, int32_t isign, int64_t batch, double eps, int32_t precision,
                          double sigma, const double *source_bounds, const double *target_bounds,
                          mnufft_plan *plan)
{
    return guarded([&] {
        if (!plan || !source_bounds || !target_bounds) {
            set_error("null argument");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        *plan = nullptr;
        if (dim < 1 || dim > mnufft::kMaxDims) {
            set_error("GPU plans take 1 to 8 dimensions, got " + std::to_string(dim));
            return MNUFFT_ERROR_PLAN;
        }
        if (batch < 1) {
            set_error("batch must be at least 1");
            return MNUFFT_ERROR_PLAN;
        }
        std::string why;
        if (!gpu_available(&why)) {
            set_error(why);
            return MNUFFT_ERROR_GPU_UNAVAILABLE;
        }
        auto p = std::make_unique<Plan>();
        p->type = 3;
        p->dim = dim;
        p->isign = isign >= 0 ? 1 : -1;
        p->batch = batch;
        p->eps = eps;
        p->sigma = sigma == 0.0 ? 2.0 : sigma;
        p->precision = precision == MNUFFT_PRECISION_F32 ? Precision::F32 : Precision::F64;
        if (!p->init_type3(source_bounds, target_bounds)) return MNUFFT_ERROR_PLAN;
        *plan = new mnufft_plan_s{std::move(p)};
        return MNUFFT_SUCCESS;
    });
}

int mnufft_type3_boxes(mnufft_plan plan, const int32_t *min_boxes, int32_t *boxes)
{
    return guarded([&] {
        if (!plan || plan->plan->type != 3) {
            set_error("target boxes are defined for type-3 plans");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        Plan &p = *plan->plan;
        if (min_boxes) {
            int k[mnufft::kMaxDims] = {1, 1, 1, 1, 1, 1, 1, 1};
            for (int a = 0; a < p.dim; ++a) {
                if (min_boxes[a] < 1 || min_boxes[a] > mnufft::kType3MaxBoxes) {
                    set_error("boxes per axis must be 1.." +
                              std::to_string(mnufft::kType3MaxBoxes));
                    return MNUFFT_ERROR_INVALID_ARGUMENT;
                }
                k[a] = min_boxes[a];
            }
            (void)hipDeviceSynchronize();  // the plan's kernels may still run
            double sb[2 * mnufft::kMaxDims], tb[2 * mnufft::kMaxDims];
            for (int a = 0; a < 2 * p.dim; ++a) {
                sb[a] = p.t3_bounds[a];
                tb[a] = p.t3_bounds[2 * mnufft::kMaxDims + a];
            }
            if (!p.init_type3(sb, tb, k)) return MNUFFT_ERROR_PLAN;
        }
        if (boxes)
            for (int a = 0; a < p.dim; ++a) boxes[a] = p.t3_boxes.k[a];
        return MNUFFT_SUCCESS;
    });
}

void mnufft_destroy(mnufft_plan plan)
{
    if (!plan) return;
    try {
        (void)hipDeviceSynchronize();
        delete plan;
    } catch (...) {
    }
}

int mnufft_set_points(mnufft_plan plan, int64_t M, const void *d_points, void *stream)
{
    return guarded([&] {
        if (!plan || (M > 0 && !d_points) || M < 0) {
            set_error("invalid plan, point count or point pointer");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        if (plan->plan->type == 3) {
            set_error("type-3 plans take their points in mnufft_execute_type3");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        if (!mnufft::point_count_ok(M, "point")) return MNUFFT_ERROR_PLAN;
        if (!plan->plan->set_points(M, d_points, (hipStream_t)stream)) return MNUFFT_ERROR_GPU;
        return MNUFFT_SUCCESS;
    });
}

int mnufft_execute(mnufft_plan plan, int64_t active_batch, const void *d_in, void *d_out,
                   void *stream)
{
    return guarded([&] {
        // a type-2 execution at zero points has no output to write
        if (!plan || (!d_out && !(plan->plan->type == 2 && plan->plan->point_count == 0)) ||
            (!d_in && plan->plan->point_count > 0)) {
            set_error("invalid plan or null buffer");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        Plan &p = *plan->plan;
        if (p.type == 3) {
            set_error("type-3 plans execute with mnufft_execute_type3");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        if (!p.points_set) {
            set_error("points must be set before executing");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        if (active_batch < 1 || active_batch > p.batch) {
            set_error("active batch " + std::to_string(active_batch) + " outside 1.." +
                      std::to_string(p.batch));
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        if (!p.execute(active_batch, d_in, d_out, (hipStream_t)stream)) return MNUFFT_ERROR_GPU;
        return MNUFFT_SUCCESS;
    });
}

int mnufft_encode(mnufft_plan plan, int64_t active_batch, int64_t M, const void *d_points,
                  const void *d_in, void *d_out, void *stream)
{
    int rc = mnufft_set_points(plan, M, d_points, stream);
    if (rc) return rc;
    return mnufft_execute(plan, active_batch, d_in, d_out, stream);
}

int mnufft_execute_type2_many(int64_t count, const mnufft_plan *plans, const int64_t *M,
                              const void *const *d_points, const void *const *d_in,
                              void *const *d_out, void *stream)
{
    return guarded([&] {
        if (count < 1 || !plans || !M || !d_points || !d_in || !d_out) {
            set_error("invalid batch arguments");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        std::vector<Plan *> ps((size_t)count);
        for (int64_t i = 0; i < count; ++i) {
            Plan *p = plans[i] ? plans[i]->plan.get() : nullptr;
            if (!p || p->type != 2 || p->precision != Precision::F32 || p->dim < 1 || p->dim > 3 ||
                p->dim != plans[0]->plan->dim) {
                set_error("a batch takes F32 type-2 plans of one dimension, 1D to 3D");
                return MNUFFT_ERROR_INVALID_ARGUMENT;
            }
            if (M[i] < 0 || !d_in[i] || (M[i] > 0 && (!d_points[i] || !d_out[i]))) {
                set_error("invalid point count or null buffer in batch");
                return MNUFFT_ERROR_INVALID_ARGUMENT;
            }
            for (int64_t k = 0; k < i; ++k)
                if (ps[(size_t)k] == p) {
                    set_error("a plan appears twice in a batch");
                    return MNUFFT_ERROR_INVALID_ARGUMENT;
                }
            ps[(size_t)i] = p;
        }
        if (!execute_type2_many(count, ps.data(), M, d_points, d_in, d_out, (hipStream_t)stream))
            return MNUFFT_ERROR_GPU;
        return MNUFFT_SUCCESS;
    });
}

int mnufft_execute_type1_many(int64_t count, const mnufft_plan *plans, const int64_t *M,
                              const void *const *d_points, const void *const *d_in,
                              void *const *d_out, void *stream)
{
    return guarded([&] {
        if (count < 1 || !plans || !M || !d_points || !d_in || !d_out) {
            set_error("invalid batch arguments");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        std::vector<Plan *> ps((size_t)count);
        for (int64_t i = 0; i < count; ++i) {
            ps[(size_t)i] = plans[i] ? plans[i]->plan.get() : nullptr;
            if (ps[(size_t)i] && !mnufft::point_count_ok(M[i], "point"))
                return MNUFFT_ERROR_PLAN;
        }
        if (!mnufft::type1_many_valid(count, ps.data(), M, d_points, d_in, d_out))
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        if (!mnufft::execute_type1_many(count, ps.data(), M, d_points, d_in, d_out,
                                        (hipStream_t)stream))
            return MNUFFT_ERROR_GPU;
        return MNUFFT_SUCCESS;
    });
}

int mnufft_execute_type3(mnufft_plan plan, int64_t active_batch, int64_t M, const void *d_sources,
                         const void *d_strengths, int64_t N, const void *d_targets, void *d_out,
                         void *stream)
{
    return guarded([&] {
        if (!plan || plan->plan->type != 3) {
            set_error("not a type-3 plan");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        Plan &p = *plan->plan;
        if (active_batch < 1 || active_batch > p.batch) {
            set_error("active batch outside 1..batch");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        if (M > mnufft::kMaxPoints || N > mnufft::kMaxPoints) {
            (void)(mnufft::point_count_ok(M, "source") && mnufft::point_count_ok(N, "target"));
            return MNUFFT_ERROR_PLAN;
        }
        if (M < 0 || N < 0 || (M > 0 && (!d_sources || !d_strengths)) ||
            (N > 0 && (!d_targets || !d_out))) {
            set_error("invalid counts or null buffers");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        if (!p.execute_type3(active_batch, M, d_sources, d_strengths, N, d_targets, d_out,
                             (hipStream_t)stream))
            return MNUFFT_ERROR_GPU;
        return MNUFFT_SUCCESS;
    });
}

int mnufft_bin_count(mnufft_plan plan, int64_t M, int64_t *nbins)
{
    return guarded([&] {
        if (!plan || !nbins || plan->plan->type != 1) {
            set_error("bin counts are defined for type-1 plans");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        *nbins = plan->plan->type1_geometry(M, nullptr).nbins;
        return MNUFFT_SUCCESS;
    });
}

int mnufft_bin_sort(mnufft_plan plan, int64_t M, const void *d_points, uint32_t *d_sorted,
                    uint32_t *d_offsets, int64_t *nbins, void *stream)
{
    return guarded([&] {
        if (!plan || plan->plan->type != 1 || (M > 0 && (!d_points || !d_sorted))) {
            set_error("bin sorting needs a type-1 plan and device buffers");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        if (!mnufft::point_count_ok(M, "point")) return MNUFFT_ERROR_PLAN;
        if (nbins) *nbins = plan->plan->type1_geometry(M, nullptr).nbins;
        if (!plan->plan->bin_sort(M, d_points, d_sorted, d_offsets, (hipStream_t)stream))
            return MNUFFT_ERROR_GPU;
        return MNUFFT_SUCCESS;
    });
}

int mnufft_bin_sort_noalloc(mnufft_plan plan, int64_t M, const void *d_points,
                            uint32_t *d_sorted, uint32_t *d_offsets, int64_t *nbins, void *stream)
{
    return guarded([&] {
        if (!plan || plan->plan->type != 1 || !d_offsets || (M > 0 && (!d_points || !d_sorted))) {
            set_error("bin sorting without allocation needs a type-1 plan, device points, a "
                      "permutation buffer and an offsets buffer");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        if (!mnufft::point_count_ok(M, "point")) return MNUFFT_ERROR_PLAN;
        const int64_t nb = plan->plan->type1_geometry(M, nullptr).nbins;
        if (nbins) *nbins = nb;
        if (M > MNUFFT_NOALLOC_MAX_POINTS || nb > MNUFFT_NOALLOC_MAX_BINS ||
            mnufft::BinSorter::external_bytes(nb, M) > mnufft::static_sort_bytes()) {
            set_error("bin sorting without allocation takes at most " +
                      std::to_string(MNUFFT_NOALLOC_MAX_POINTS) + " points and " +
                      std::to_string(MNUFFT_NOALLOC_MAX_BINS) + " bins (" + std::to_string(M) +
                      " points in " + std::to_string(nb) + " bins asked)");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        if (!plan->plan->bin_sort_static(M, d_points, d_sorted, d_offsets, (hipStream_t)stream))
            return MNUFFT_ERROR_GPU;
        return MNUFFT_SUCCESS;
    });
}

int mnufft_points_permutation(mnufft_plan plan, uint32_t *d_sorted, uint32_t *d_offsets,
                              int64_t *nbins, int64_t *bin_side, void *stream)
{
    return guarded([&] {
        if (!plan || plan->plan->type != 1) {
            set_error("the point permutation is defined for type-1 plans");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        Plan &p = *plan->plan;
        if (!p.points_set || p.binned_dense != 0u) {
            set_error(p.points_set ? "dense bins keep their records apart from the bin order"
                                   : "no points are binned");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        if (nbins) *nbins = p.gp.nbins;
        if (bin_side)
            for (int a = 0; a < std::max(3, p.dim); ++a)
                bin_side[a] = a < p.dim ? ((int64_t)1 << p.gp.ax[a].shift) : 1;
        hipStream_t st = (hipStream_t)stream;
        if (p.point_count > 0 && d_sorted &&
            !mnufft::hip_ok(hipMemcpyAsync(d_sorted, p.rec_perm, sizeof(uint32_t) * p.point_count,
                                           hipMemcpyDeviceToDevice, st),
                            "permutation copy"))
            return MNUFFT_ERROR_GPU;
        if (d_offsets &&
            !mnufft::hip_ok(hipMemcpyAsync(d_offsets, p.sorter.offsets(),
                                           sizeof(uint32_t) * (p.gp.nbins + 1),
                                           hipMemcpyDeviceToDevice, st),
                            "offsets copy"))
            return MNUFFT_ERROR_GPU;
        return MNUFFT_SUCCESS;
    });
}

int mnufft_plan_info(mnufft_plan plan, int64_t *info, int32_t n)
{
    if (!plan || !info) {
        set_error("null argument");
        return MNUFFT_ERROR_INVALID_ARGUMENT;
    }
    const Plan &p = *plan->plan;
    const Plan &q = p.type == 3 && p.outer ? *p.outer : p;
    int64_t v[16] = {0};
    v[0] = q.kernel.width;
    for (int a = 0; a < 3; ++a) {
        v[1 + a] = a < q.dim ? q.fine[a] : 1;
        v[4 + a] = a < q.dim ? ((int64_t)1 << q.gp.ax[a].shift) : 1;
        v[9 + a] = a < q.dim ? q.spread_params().block[a] : 1;
        v[13 + a] = a < p.dim ? p.type3_outer[a] : 1;
    }
    v[7] = q.gp.nbins;
    v[8] = q.geometry_index;
    v[12] = q.spread_params().run;
    for (int i = 0; i < n && i < 16; ++i) info[i] = v[i];
    return MNUFFT_SUCCESS;
}

double mnufft_plan_beta(mnufft_plan plan) { return plan ? plan->plan->kernel.beta : 0.0; }

int mnufft_stage_times(mnufft_plan plan, double *ms, int32_t n)
{
    return guarded([&] {
        if (!plan || !ms) {
            set_error("null argument");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        return plan->plan->stage_times(ms, n, nullptr) ? MNUFFT_SUCCESS : MNUFFT_ERROR_GPU;
    });
}

int mnufft_timing_history(mnufft_plan plan, int32_t on)
{
    if (!plan) {
        set_error("null plan");
        return MNUFFT_ERROR_INVALID_ARGUMENT;
    }
    plan->plan->history = on != 0;
    plan->plan->slots_filled = 0;
    return MNUFFT_SUCCESS;
}

int mnufft_set_spread_method(mnufft_plan plan, int32_t method)
{
    if (!plan) {
        set_error("null plan");
        return MNUFFT_ERROR_INVALID_ARGUMENT;
    }
    Plan &p = *plan->plan;
    Plan &q = p.type == 3 && p.outer ? *p.outer : p;
    if (method < MNUFFT_SPREAD_PLAIN || method > MNUFFT_SPREAD_MFMA_BF16X3) {
        set_error("unknown spread method");
        return MNUFFT_ERROR_INVALID_ARGUMENT;
    }
    if (method != MNUFFT_SPREAD_PLAIN && q.type == 1 && q.dim == 2) {
        // The 2D matrix-core tiles (32 x 32 cells) ended in a GPU memory
        // fault on a 1024^2-mode, 5e7-point plan whose cause is not found;
        // they stay switched off until it is.
        set_error("the 2D matrix-core spread is switched off (an open GPU memory fault at "
                  "5e7 points); 2D plans spread in plain FP32");
        return MNUFFT_ERROR_INVALID_ARGUMENT;
    }
    if (method != MNUFFT_SPREAD_PLAIN &&
        (q.type != 1 || q.precision != Precision::F32 || !(q.dim == 3 && q.sweep))) {
        set_error("matrix-core spreading needs an FP32 3D type-1 plan with at least 16 + width "
                  "fine cells per axis and a width <= 8");
        return MNUFFT_ERROR_INVALID_ARGUMENT;
    }
    q.spread_method = (mnufft::SpreadMethod)method;
    return MNUFFT_SUCCESS;
}

int mnufft_autotune(mnufft_plan plan, int64_t active_batch, int64_t M, const void *d_points,
                    const void *d_in, void *d_out, int32_t repeats, int64_t *choice, void *stream)
{
    return guarded([&] {
        if (!plan || plan->plan->type != 1 || M < 0 || (M > 0 && (!d_points || !d_in)) ||
            !d_out || repeats < 1 || repeats > 100) {
            set_error("autotune takes a type-1 plan, its points, strengths, an output and "
                      "1..100 repeats");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        Plan &p = *plan->plan;
        if (active_batch < 1 || active_batch > p.batch) {
            set_error("active batch outside 1..batch");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        if (!mnufft::point_count_ok(M, "point")) return MNUFFT_ERROR_PLAN;
        // candidates: the plain spread, and the matrix-core methods whose
        // operand error stays below the plan's tolerance, each 3D one with
        // dense bins on and off
        struct Cand {
            mnufft::SpreadMethod method;
            uint32_t dense;
        };
        std::vector<Cand> cands{{mnufft::SpreadMethod::Plain, 0u}};
        // (the 2D matrix-core tiles are switched off: an open GPU fault)
        const bool mfma_ok = p.precision == Precision::F32 && p.dim == 3 && p.sweep;
        if (mfma_ok) {
            std::vector<mnufft::SpreadMethod> ms{mnufft::SpreadMethod::MfmaF16x3};
            if (p.kernel.eps >= 1e-4) ms.push_back(mnufft::SpreadMethod::MfmaBf16x3);
            if (p.kernel.eps >= 1e-3) ms.push_back(mnufft::SpreadMethod::MfmaF16);
            for (mnufft::SpreadMethod m : ms) {
                cands.push_back({m, 0u});
                if (p.dim == 3) cands.push_back({m, 32u});
            }
        }
        const mnufft::SpreadMethod keep_method = p.spread_method;
        const uint32_t keep_dense = p.dense.threshold;
        hipStream_t st = (hipStream_t)stream;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (!mnufft::hip_ok(hipEventCreate(&e0), "autotune event") ||
            !mnufft::hip_ok(hipEventCreate(&e1), "autotune event"))
            return MNUFFT_ERROR_GPU;
        double best = -1.0;
        size_t best_i = 0;
        bool ok = true;
        std::vector<float> times;
        for (size_t i = 0; i < cands.size() && ok; ++i) {
            p.spread_method = cands[i].method;
            p.dense.threshold = cands[i].dense;
            // one untimed encode (sizes the scratch), then `repeats` timed
            ok = p.set_points(M, d_points, st) && p.execute(active_batch, d_in, d_out, st);
            ok = ok && mnufft::hip_ok(hipEventRecord(e0, st), "autotune event");
            for (int r = 0; r < repeats && ok; ++r)
                ok = p.set_points(M, d_points, st) && p.execute(active_batch, d_in, d_out, st);
            ok = ok && mnufft::hip_ok(hipEventRecord(e1, st), "autotune event") &&
                 mnufft::hip_ok(hipEventSynchronize(e1), "autotune wait");
            float ms = 0.0f;
            ok = ok && mnufft::hip_ok(hipEventElapsedTime(&ms, e0, e1), "autotune timing");
            // dense bins are kept only when they gain at least 3% over the
            // same method without them (they cost set-up work and scratch)
            const bool dense_gain =
                cands[i].dense == 0u || i == 0 || cands[i - 1].method != cands[i].method ||
                ms < 0.97 * times.back();
            times.push_back(ms);
            if (ok && dense_gain && (best < 0.0 || ms < best)) {
                best = ms;
                best_i = i;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        if (!ok) {
            p.spread_method = keep_method;
            p.dense.threshold = keep_dense;
            return MNUFFT_ERROR_GPU;
        }
        p.spread_method = cands[best_i].method;
        p.dense.threshold = cands[best_i].dense;
        int64_t tried = (int64_t)cands.size();
        // Then, for the method kept, the bins and tiles: the plain 2D/3D
        // spread on the output-stationary blocks of the reference layout
        // instead of the row tiles; coarser bins of the row geometry along axis 0 (2^k
        // cells); the heavy-tile split (light limit).  Each is kept when it
        // gains at least 3% over the best so far.
        struct Knob {
            bool plain_rows;
            int64_t side0;
            uint32_t light;
        };
        const Knob base{p.plain_rows, p.rows_side0, p.spreader.light_limit};
        std::vector<Knob> knobs;
        const bool rows = p.sweep || (p.rows2 && p.spread_method == mnufft::SpreadMethod::Plain);
        if (rows && p.spread_method == mnufft::SpreadMethod::Plain)
            knobs.push_back({!base.plain_rows, base.side0, base.light});
        if (rows)
            for (int64_t s0 : {4, 8}) knobs.push_back({base.plain_rows, s0, base.light});
        for (uint32_t lim : {1024u, 16384u}) knobs.push_back({base.plain_rows, base.side0, lim});
        Knob keep = base;
        if (!mnufft::hip_ok(hipEventCreate(&e0), "autotune event") ||
            !mnufft::hip_ok(hipEventCreate(&e1), "autotune event"))
            return MNUFFT_ERROR_GPU;
        for (const Knob &k : knobs) {
            // each knob on top of the best so far
            Knob cand = keep;
            if (k.plain_rows != base.plain_rows) cand.plain_rows = k.plain_rows;
            if (k.side0 != base.side0) cand.side0 = k.side0;
            if (k.light != base.light) cand.light = k.light;
            p.plain_rows = cand.plain_rows;
            p.rows_side0 = cand.side0;
            p.spreader.light_limit = cand.light;
            ok = p.set_points(M, d_points, st) && p.execute(active_batch, d_in, d_out, st);
            ok = ok && mnufft::hip_ok(hipEventRecord(e0, st), "autotune event");
            for (int r = 0; r < repeats && ok; ++r)
                ok = p.set_points(M, d_points, st) && p.execute(active_batch, d_in, d_out, st);
            ok = ok && mnufft::hip_ok(hipEventRecord(e1, st), "autotune event") &&
                 mnufft::hip_ok(hipEventSynchronize(e1), "autotune wait");
            float ms = 0.0f;
            ok = ok && mnufft::hip_ok(hipEventElapsedTime(&ms, e0, e1), "autotune timing");
            if (!ok) break;
            ++tried;
            if (ms < 0.97 * best) {
                best = ms;
                keep = cand;
            }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        p.plain_rows = keep.plain_rows;
        p.rows_side0 = keep.side0;
        p.spreader.light_limit = keep.light;
        if (!ok) return MNUFFT_ERROR_GPU;
        // without dense bins their scratch is given back (it cannot be
        // needed before the threshold changes again)
        if (cands[best_i].dense == 0u) p.dense.release();
        if (choice) {
            choice[0] = (int64_t)cands[best_i].method;
            choice[1] = (int64_t)cands[best_i].dense;
            choice[2] = (int64_t)(best / repeats * 1e6);  // ns per encode
            choice[3] = tried;
        }
        return MNUFFT_SUCCESS;
    });
}

int mnufft_set_point_order(mnufft_plan plan, int32_t order)
{
    if (!plan) {
        set_error("null plan");
        return MNUFFT_ERROR_INVALID_ARGUMENT;
    }
    Plan &p = *plan->plan;
    Plan &q = p.type == 3 && p.outer ? *p.outer : p;
    if (order != MNUFFT_POINT_ORDER_INDEX && order != MNUFFT_POINT_ORDER_COORDINATES) {
        set_error("unknown point order");
        return MNUFFT_ERROR_INVALID_ARGUMENT;
    }
    if (q.type != 1) {
        set_error("the point order applies to type-1 and type-3 plans");
        return MNUFFT_ERROR_INVALID_ARGUMENT;
    }
    q.point_order = order == MNUFFT_POINT_ORDER_INDEX ? mnufft::BinOrder::Index
                                                      : mnufft::BinOrder::Canonical;
    return MNUFFT_SUCCESS;
}

int mnufft_set_spread_tuning(mnufft_plan plan, int64_t light_limit, int64_t min_part,
                             int64_t part_budget_bytes)
{
    if (!plan || light_limit < 0 || min_part < 0 || part_budget_bytes < 0 ||
        light_limit > 0x7fffffff || min_part > 0x7fffffff) {
        set_error("invalid spread tuning");
        return MNUFFT_ERROR_INVALID_ARGUMENT;
    }
    Plan &p = *plan->plan;
    Plan &q = p.type == 3 && p.outer ? *p.outer : p;
    if (light_limit) q.spreader.light_limit = (uint32_t)light_limit;
    if (min_part) q.spreader.min_part = (uint32_t)min_part;
    if (part_budget_bytes) q.spreader.budget_bytes = part_budget_bytes;
    return MNUFFT_SUCCESS;
}

int mnufft_set_bin_tuning(mnufft_plan plan, int64_t rows_side0, int32_t plain_rows)
{
    if (!plan) {
        set_error("null plan");
        return MNUFFT_ERROR_INVALID_ARGUMENT;
    }
    if (rows_side0 < 0 || rows_side0 > 64 || (rows_side0 & (rows_side0 - 1)) != 0) {
        set_error("the axis-0 bin side is 0 (automatic) or a power of two up to 64");
        return MNUFFT_ERROR_INVALID_ARGUMENT;
    }
    Plan &p = *plan->plan;
    Plan &q = p.type == 3 && p.outer ? *p.outer : p;
    q.rows_side0 = rows_side0;
    q.plain_rows = plain_rows != 0;
    return MNUFFT_SUCCESS;
}

int mnufft_set_fft_engine(mnufft_plan plan, int32_t engine)
{
    if (!plan) {
        set_error("null plan");
        return MNUFFT_ERROR_INVALID_ARGUMENT;
    }
    if (engine != MNUFFT_FFT_LINES && engine != MNUFFT_FFT_ROCFFT) {
        set_error("the FFT engine is MNUFFT_FFT_LINES (0) or MNUFFT_FFT_ROCFFT (1)");
        return MNUFFT_ERROR_INVALID_ARGUMENT;
    }
    Plan &p = *plan->plan;
    Plan &q = p.type == 3 && p.inner ? *p.inner : p;
    q.fft_engine = engine;
    return MNUFFT_SUCCESS;
}

int mnufft_fft_engine(mnufft_plan plan, int32_t *engine)
{
    if (!plan || !engine) {
        set_error("null argument");
        return MNUFFT_ERROR_INVALID_ARGUMENT;
    }
    const Plan &p = *plan->plan;
    const Plan &q = p.type == 3 && p.inner ? *p.inner : p;
    *engine = q.use_line_fft() ? MNUFFT_FFT_LINES : MNUFFT_FFT_ROCFFT;
    return MNUFFT_SUCCESS;
}

int mnufft_plan_tuning(mnufft_plan plan, int64_t *vals, int32_t n)
{
    if (!plan || !vals) {
        set_error("null argument");
        return MNUFFT_ERROR_INVALID_ARGUMENT;
    }
    const Plan &p = *plan->plan;
    const Plan &q = p.type == 3 && p.outer ? *p.outer : p;
    const int64_t v[6] = {(int64_t)q.spread_method, (int64_t)q.dense.threshold, q.rows_side0,
                          q.plain_rows ? 1 : 0, (int64_t)q.spreader.light_limit,
                          (int64_t)q.spreader.min_part};
    for (int i = 0; i < n && i < 6; ++i) vals[i] = v[i];
    return MNUFFT_SUCCESS;
}

int mnufft_set_dense_bins(mnufft_plan plan, int64_t min_records)
{
    if (!plan || min_records < 0 || min_records > 0x7fffffff) {
        set_error("invalid dense-bin threshold");
        return MNUFFT_ERROR_INVALID_ARGUMENT;
    }
    Plan &p = *plan->plan;
    Plan &q = p.type == 3 && p.outer ? *p.outer : p;
    q.dense.threshold = (uint32_t)min_records;
    return MNUFFT_SUCCESS;
}

int mnufft_direct_f64(int32_t kind, int32_t dim, const int64_t *n_modes, int32_t isign,
                      int32_t mode_order, int64_t M, const double *d_x, const double *d_in,
                      int64_t nout, const int64_t *d_sel, const double *d_s, double *d_out,
                      void *stream)
{
    return guarded([&] {
        if (kind < 1 || kind > 3 || dim < 1 || dim > mnufft::kMaxDims || !d_out ||
            (kind != 3 && !n_modes)) {
            set_error("invalid direct-sum arguments");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        mnufft::DirectParams dp{};
        dp.mode_total = 1;
        for (int a = 0; a < mnufft::kMaxDims; ++a) {
            dp.n_modes[a] = (kind != 3 && a < dim) ? n_modes[a] : 1;
            dp.mode_total *= dp.n_modes[a];
        }
        dp.M = M;
        dp.nout = nout;
        dp.centered = mode_order == 0;
        dp.sign = isign >= 0 ? 1.0 : -1.0;
        if (nout == 0) return MNUFFT_SUCCESS;
        return mnufft::direct_sum(kind, dim, dp, d_x, d_in, d_sel, d_s, d_out,
                                  (hipStream_t)stream)
                   ? MNUFFT_SUCCESS
                   : MNUFFT_ERROR_GPU;
    });
}

int mnufft_kernel_params(double eps, double sigma, int32_t *width, double *beta)
{
    mnufft::EsKernel k;
    if (!width || !beta) {
        set_error("null argument");
        return MNUFFT_ERROR_INVALID_ARGUMENT;
    }
    if (!mnufft::es_kernel_for_tolerance(eps, sigma == 0.0 ? 2.0 : sigma, &k))
        return MNUFFT_ERROR_PLAN;
    *width = k.width;
    *beta = k.beta;
    return MNUFFT_SUCCESS;
}

int mnufft_tolerance_check(int32_t dim, const int64_t *n_modes, double eps, double sigma,
                           int32_t precision)
{
    return guarded([&] {
        if (dim < 1 || dim > 8 || !n_modes) {
            set_error("tolerance check takes 1 to 8 axes and their mode counts");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        const double sg = sigma == 0.0 ? 2.0 : sigma;
        mnufft::EsKernel k;
        if (!mnufft::es_kernel_for_tolerance(eps, sg, &k)) return MNUFFT_ERROR_PLAN;
        std::vector<int64_t> nm((size_t)dim);
        std::vector<std::vector<double>> coeffs;
        for (int a = 0; a < dim; ++a) {
            if (n_modes[a] < 1) {
                set_error("mode count along axis " + std::to_string(a) + " must be positive");
                return MNUFFT_ERROR_INVALID_ARGUMENT;
            }
            nm[(size_t)a] = n_modes[a];
            const int64_t n = mnufft::select_fine_grid_size(n_modes[a], sg, k.width);
            if (n < 0) {
                set_error("no supported fine-grid length for axis " + std::to_string(a));
                return MNUFFT_ERROR_PLAN;
            }
            coeffs.push_back(mnufft::centered_fourier_coefficients(k, n));
        }
        const Precision pr = precision == MNUFFT_PRECISION_F32    ? Precision::F32
                             : precision == MNUFFT_PRECISION_DF64 ? Precision::Df64
                                                                  : Precision::F64;
        return mnufft::validate_f32_rounding(nm, eps, sg, pr, coeffs) ? MNUFFT_SUCCESS
                                                                     : MNUFFT_ERROR_PLAN;
    });
}

int64_t mnufft_fine_grid_size(int64_t n_modes, double sigma, int32_t width)
{
    return mnufft::select_fine_grid_size(n_modes, sigma == 0.0 ? 2.0 : sigma, width);
}

int mnufft_layout_query(int32_t dim, const int64_t *fine, int32_t width, int32_t precision,
                        int64_t point_count, int64_t *out)
{
    return guarded([&] {
        if (dim < 1 || dim > mnufft::kMaxDims || !fine || !out) {
            set_error("invalid layout query");
            return MNUFFT_ERROR_INVALID_ARGUMENT;
        }
        mnufft::SpreadLayout L = mnufft::build_spread_layout(
            dim, fine, width, precision == MNUFFT_PRECISION_F32 ? Precision::F32 : Precision::F64);
        for (int i = 0; i < 128; ++i) out[i] = 0;
        out[0] = L.run;
        out[1] = L.workgroup_size;
        out[2] = L.row_batch;
        out[3] = (int64_t)L.geometries.size();
        out[4] = L.choose_geometry(point_count);
        for (int a = 0; a < mnufft::kMaxDims; ++a) {
            out[8 + a] = a < dim ? L.block[a] : 1;
            out[16 + a] = a < dim ? L.blocks[a] : 1;
        }
        for (size_t g = 0; g < L.geometries.size() && g < 12; ++g)
            for (int a = 0; a < mnufft::kMaxDims; ++a)
                out[32 + 8 * g + a] = a < dim ? L.geometries[g].side[a] : 1;
        return MNUFFT_SUCCESS;
    });
}

}  // extern "C"

// ======================================================================
// Host-array API (reference C ABI).
// ======================================================================
namespace {

struct HostPlan {
    int type = 1;
    int dim = 1;
    int64_t n_modes[mnufft::kMaxDims] = {1, 1, 1, 1, 1, 1, 1, 1};
    int isign = 1;
    int64_t ntrans = 1;
    double eps = 1e-6;
    double sigma = 2.0;
    int mode_order = 0;
    Precision precision = Precision::F64;
    bool float_api = false;
    std::unique_ptr<Plan> plan;
    // device copies
    void *d_points = nullptr, *d_targets = nullptr, *d_in = nullptr, *d_out = nullptr;
    size_t points_bytes = 0, targets_bytes = 0, in_bytes = 0, out_bytes = 0;
    int64_t M = 0, N = 0;
    bool points_set = false;
    ~HostPlan()
    {
        (void)hipDeviceSynchronize();
        for (void *p : {d_points, d_targets, d_in, d_out})
            if (p) (void)hipFree(p);
    }
    size_t real() const { return precision == Precision::F32 ? 4 : 8; }
};

bool grow(void **p, size_t *cap, size_t need)
{
    if (need <= *cap && *p) return true;
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    *cap = 0;
    if (!mnufft::hip_ok(hipMalloc(p, need ? need : 16), "device buffer")) return false;
    *cap = need;
    return true;
}

int hp_makeplan(bool float_api, int32_t type, int32_t dim, const int64_t *n_modes, int32_t isign,
                int64_t ntrans, double eps, HostPlan **out, const wgpu_nufft_opts *opts)
{
    if (!out) {
        set_error("plan output pointer is null");
        return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
    }
    *out = nullptr;
    wgpu_nufft_opts o;
    std::memset(&o, 0, sizeof(o));
    if (opts) o = *opts;
    if (o.backend == WGPU_NUFFT_BACKEND_CPU) {
        set_error("this engine has no CPU backend; use BACKEND_AUTO or BACKEND_GPU");
        return WGPU_NUFFT_ERROR_GPU_UNAVAILABLE;
    }
    if (o.backend < 0 || o.backend > 2 || o.precision < 0 || o.precision > 3 ||
        (o.mode_order != 0 && o.mode_order != 1) || o.threads < 0 || !(o.sigma >= 0.0)) {
        set_error("invalid options");
        return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
    }
    if (o.adapter_index > 1 || (o.adapter_index < 0)) {
        set_error("adapter selection matches no HIP device");
        return WGPU_NUFFT_ERROR_GPU_UNAVAILABLE;
    }
    if (type < 1 || type > 3) {
        set_error("type must be 1, 2 or 3");
        return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
    }
    if (dim < 1 || dim > mnufft::kMaxDims) {
        set_error("dim must be between 1 and 8, not " + std::to_string(dim));
        return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
    }
    if (ntrans < 1) {
        set_error("ntrans must be at least 1");
        return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
    }
    if (type != 3) {
        if (!n_modes) {
            set_error("n_modes is null");
            return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
        }
        for (int a = 0; a < dim; ++a)
            if (n_modes[a] < 1) {
                set_error("mode count along axis " + std::to_string(a) + " must be positive");
                return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
            }
    }
    std::string why;
    if (!gpu_available(&why)) {
        set_error(why);
        return WGPU_NUFFT_ERROR_GPU_UNAVAILABLE;
    }
    auto hp = std::make_unique<HostPlan>();
    hp->type = type;
    hp->dim = dim;
    hp->isign = isign >= 0 ? 1 : -1;
    hp->ntrans = ntrans;
    hp->eps = eps;
    hp->sigma = o.sigma == 0.0 ? 2.0 : o.sigma;
    hp->mode_order = o.mode_order;
    hp->float_api = float_api;
    hp->precision = (float_api || o.precision == WGPU_NUFFT_PRECISION_F32) ? Precision::F32
                                                                          : Precision::F64;
    if (type != 3) {
        if (!n_modes) {
            set_error("n_modes is null");
            return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
        }
        for (int a = 0; a < dim; ++a) hp->n_modes[a] = n_modes[a];
        int prec = hp->precision == Precision::F32 ? MNUFFT_PRECISION_F32 : MNUFFT_PRECISION_F64;
        int rc = make_plan(type, dim, n_modes, isign, ntrans, eps, prec, o.mode_order, hp->sigma,
                           &hp->plan);
        if (rc) return rc;
    } else {
        mnufft::EsKernel k;
        if (!mnufft::es_kernel_for_tolerance(eps, hp->sigma, &k)) return WGPU_NUFFT_ERROR_PLAN;
    }
    *out = hp.release();
    return WGPU_NUFFT_SUCCESS;
}

// Interleaves per-axis host arrays (or takes point-major ones) into the
// plan's arithmetic type, validating the coordinate range of types 1 and 2.
// An F32 plan sees coordinates rounded to float, so they are rounded before
// the checks, and lo/hi (type-3 bounds, per axis) come from the rounded
// values, as the reference does (capi/src/plan.rs:486-499).
template <typename S>
bool pack_points(const HostPlan &hp, int64_t M, const S *const *axes, const S *nd,
                 std::vector<unsigned char> *out, bool check_range, const char *what,
                 double *lo = nullptr, double *hi = nullptr)
{
    int d = hp.dim;
    out->resize((size_t)M * d * hp.real());
    for (int a = 0; a < d && lo && hi; ++a) lo[a] = hi[a] = 0.0;
    for (int64_t j = 0; j < M; ++j)
        for (int a = 0; a < d; ++a) {
            double v = nd ? (double)nd[j * d + a] : (double)axes[a][j];
            if (hp.precision == Precision::F32) v = (double)(float)v;
            if (!std::isfinite(v) || (check_range && std::fabs(v) > kMaxCoordinate)) {
                set_error(std::string(what) + " " + std::to_string(j) + " axis " +
                          std::to_string(a) + " is " + std::to_string(v) +
                          (check_range ? ", outside [-3*pi, 3*pi]" : ", not finite"));
                return false;
            }
            if (lo && hi) {
                lo[a] = j == 0 ? v : std::min(lo[a], v);
                hi[a] = j == 0 ? v : std::max(hi[a], v);
            }
            if (hp.precision == Precision::F32) {
                const float f = (float)v;
                std::memcpy(out->data() + ((size_t)j * d + a) * 4, &f, 4);
            } else {
                std::memcpy(out->data() + ((size_t)j * d + a) * 8, &v, 8);
            }
        }
    return true;
}

template <typename S>
int hp_setpts(HostPlan *hp, int64_t M, const S *const *axes, const S *nd, int64_t N,
              const S *const *taxes, const S *tnd)
{
    if (!hp) {
        set_error("plan is null");
        return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
    }
    if ((sizeof(S) == 4) != hp->float_api) {
        set_error(std::string("the plan takes ") + (hp->float_api ? "float" : "double") +
                  " arrays: this call has the other array type");
        return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
    }
    if (M > mnufft::kMaxPoints || N > mnufft::kMaxPoints) {
        (void)(mnufft::point_count_ok(M, "point") && mnufft::point_count_ok(N, "target"));
        return WGPU_NUFFT_ERROR_PLAN;
    }
    if (!nd && hp->dim > 3) {
        // (capi/src/lib.rs:190-194)
        set_error("plans of more than three dimensions take their points through setpts_nd");
        return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
    }
    if (M < 0 || (M > 0 && !nd && !axes[0]) || (hp->dim > 1 && M > 0 && !nd && !axes[1]) ||
        (hp->dim > 2 && M > 0 && !nd && !axes[2])) {
        set_error("point count or coordinate arrays invalid");
        return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
    }
    std::vector<unsigned char> pts;
    double slo[mnufft::kMaxDims], shi[mnufft::kMaxDims];
    if (!pack_points<S>(*hp, M, axes, nd, &pts, hp->type != 3, "point", slo, shi))
        return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
    if (!grow(&hp->d_points, &hp->points_bytes, pts.size())) return WGPU_NUFFT_ERROR_GPU;
    if (!pts.empty() &&
        !mnufft::hip_ok(hipMemcpy(hp->d_points, pts.data(), pts.size(), hipMemcpyHostToDevice),
                        "point upload"))
        return WGPU_NUFFT_ERROR_GPU;
    hp->M = M;
    if (hp->type == 3) {
        if (N < 0 || (N > 0 && !tnd && (!taxes[0] || (hp->dim > 1 && !taxes[1]) ||
                                         (hp->dim > 2 && !taxes[2])))) {
            set_error("target count or target arrays invalid");
            return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
        }
        std::vector<unsigned char> tg;
        double tlo[mnufft::kMaxDims], thi[mnufft::kMaxDims];
        if (!pack_points<S>(*hp, N, taxes, tnd, &tg, false, "target", tlo, thi))
            return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
        if (!grow(&hp->d_targets, &hp->targets_bytes, tg.size())) return WGPU_NUFFT_ERROR_GPU;
        if (!tg.empty() &&
            !mnufft::hip_ok(hipMemcpy(hp->d_targets, tg.data(), tg.size(), hipMemcpyHostToDevice),
                            "target upload"))
            return WGPU_NUFFT_ERROR_GPU;
        hp->N = N;
        // bounds from the points the plan computes with (rounded for F32),
        // as the plan's caller guarantee
        double sb[2 * mnufft::kMaxDims], tb[2 * mnufft::kMaxDims];
        for (int a = 0; a < hp->dim; ++a) {
            sb[2 * a] = M > 0 ? slo[a] : 0.0;
            sb[2 * a + 1] = M > 0 ? shi[a] : 0.0;
            tb[2 * a] = N > 0 ? tlo[a] : 0.0;
            tb[2 * a + 1] = N > 0 ? thi[a] : 0.0;
        }
        auto p = std::make_unique<Plan>();
        p->type = 3;
        p->dim = hp->dim;
        p->isign = hp->isign;
        p->batch = hp->ntrans;
        p->eps = hp->eps;
        p->sigma = hp->sigma;
        p->precision = hp->precision;
        if (!p->init_type3(sb, tb)) return WGPU_NUFFT_ERROR_PLAN;
        hp->plan = std::move(p);
    } else {
        if (!hp->plan->set_points(M, hp->d_points, nullptr)) return WGPU_NUFFT_ERROR_GPU;
    }
    hp->points_set = true;
    return WGPU_NUFFT_SUCCESS;
}

template <typename S>
int hp_execute(HostPlan *hp, S *c, S *f)
{
    if (!hp) {
        set_error("plan is null");
        return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
    }
    if ((sizeof(S) == 4) != hp->float_api) {
        set_error(std::string("the plan takes ") + (hp->float_api ? "float" : "double") +
                  " arrays: this call has the other array type");
        return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
    }
    if (!hp->points_set) {
        set_error("set the points (setpts) before execute");
        return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
    }
    int64_t modes = 1;
    for (int a = 0; a < hp->dim; ++a) modes *= hp->n_modes[a];
    int64_t nin, nout;
    const S *src;
    S *dst;
    if (hp->type == 1) {
        nin = hp->M * hp->ntrans;
        nout = modes * hp->ntrans;
        src = c;
        dst = f;
    } else if (hp->type == 2) {
        nin = modes * hp->ntrans;
        nout = hp->M * hp->ntrans;
        src = f;
        dst = c;
    } else {
        nin = hp->M * hp->ntrans;
        nout = hp->N * hp->ntrans;
        src = c;
        dst = f;
    }
    if ((nin > 0 && !src) || (nout > 0 && !dst)) {
        set_error("input or output array is null");
        return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
    }
    size_t rb = hp->real();
    // widen or narrow to the plan's arithmetic
    std::vector<unsigned char> in((size_t)nin * 2 * rb);
    for (int64_t i = 0; i < 2 * nin; ++i) {
        if (!std::isfinite((double)src[i])) {
            set_error("input value " + std::to_string(i / 2) + " is not finite");
            return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
        }
        if (rb == 4) {
            float v = (float)src[i];
            std::memcpy(in.data() + 4 * i, &v, 4);
        } else {
            double v = (double)src[i];
            std::memcpy(in.data() + 8 * i, &v, 8);
        }
    }
    if (!grow(&hp->d_in, &hp->in_bytes, in.size()) ||
        !grow(&hp->d_out, &hp->out_bytes, (size_t)nout * 2 * rb))
        return WGPU_NUFFT_ERROR_GPU;
    if (!in.empty() &&
        !mnufft::hip_ok(hipMemcpy(hp->d_in, in.data(), in.size(), hipMemcpyHostToDevice),
                        "input upload"))
        return WGPU_NUFFT_ERROR_GPU;
    bool ok;
    if (hp->type == 3)
        ok = hp->plan->execute_type3(hp->ntrans, hp->M, hp->d_points, hp->d_in, hp->N,
                                     hp->d_targets, hp->d_out, nullptr);
    else
        ok = hp->plan->execute(hp->ntrans, hp->d_in, hp->d_out, nullptr);
    if (!ok) return WGPU_NUFFT_ERROR_GPU;
    if (hp->type == 2 && hp->M == 0) return WGPU_NUFFT_SUCCESS;
    std::vector<unsigned char> out((size_t)nout * 2 * rb);
    if (!out.empty() &&
        !mnufft::hip_ok(hipMemcpy(out.data(), hp->d_out, out.size(), hipMemcpyDeviceToHost),
                        "output download"))
        return WGPU_NUFFT_ERROR_GPU;
    for (int64_t i = 0; i < 2 * nout; ++i) {
        if (rb == 4) {
            float v;
            std::memcpy(&v, out.data() + 4 * i, 4);
            dst[i] = (S)v;
        } else {
            double v;
            std::memcpy(&v, out.data() + 8 * i, 8);
            dst[i] = (S)v;
        }
    }
    return WGPU_NUFFT_SUCCESS;
}

std::mutex g_name_mutex;
std::string g_gpu_name;

// Plans the one-call transforms keep for reuse (capi/src/cache.rs:10-51):
// up to kOneCallPlans, most recently used last, keyed by everything plan
// creation reads (API precision, type, rank, modes, sign, tolerance, the
// options' bytes; no options = the defaults).  A loop of one-call
// transforms of the same sizes then pays plan creation (rocFFT plans,
// amplitude tables, scratch) once.  Each kept plan runs one call at a time;
// wgpu_nufft_shutdown drops them.
constexpr size_t kOneCallPlans = 4;

struct OneCallKey {
    bool float_api;
    int type, dim, isign;
    int64_t nm[3];
    double eps;
    unsigned char opts[sizeof(wgpu_nufft_opts)];
    bool operator==(const OneCallKey &o) const
    {
        return float_api == o.float_api && type == o.type && dim == o.dim && isign == o.isign &&
               std::memcmp(nm, o.nm, sizeof(nm)) == 0 &&
               std::memcmp(&eps, &o.eps, sizeof(eps)) == 0 &&
               std::memcmp(opts, o.opts, sizeof(opts)) == 0;
    }
};

struct KeptPlan {
    std::mutex mu;
    std::unique_ptr<HostPlan> hp;
};

std::mutex g_kept_mu;
// never destroyed: kept plans hold device memory, which must not be freed
// from static destructors after the HIP runtime has shut down (the process's
// exit reclaims it)
auto &g_kept = *new std::vector<std::pair<OneCallKey, std::shared_ptr<KeptPlan>>>();

int kept_plan_for(const OneCallKey &key, const int64_t *nm, const wgpu_nufft_opts *opts,
                  std::shared_ptr<KeptPlan> *out)
{
    {
        std::lock_guard<std::mutex> lock(g_kept_mu);
        for (size_t i = 0; i < g_kept.size(); ++i) {
            if (g_kept[i].first == key) {
                auto entry = g_kept[i];
                g_kept.erase(g_kept.begin() + (std::ptrdiff_t)i);
                g_kept.push_back(entry);
                *out = entry.second;
                return WGPU_NUFFT_SUCCESS;
            }
        }
    }
    // created outside the lock, so other threads' calls are not held up
    HostPlan *hp = nullptr;
    int rc = hp_makeplan(key.float_api, key.type, key.dim, nm, key.isign, 1, key.eps, &hp, opts);
    if (rc) return rc;
    auto kept = std::make_shared<KeptPlan>();
    kept->hp.reset(hp);
    std::lock_guard<std::mutex> lock(g_kept_mu);
    if (g_kept.size() == kOneCallPlans) g_kept.erase(g_kept.begin());
    g_kept.emplace_back(key, kept);
    *out = kept;
    return WGPU_NUFFT_SUCCESS;
}

void drop_kept_plans()
{
    std::vector<std::pair<OneCallKey, std::shared_ptr<KeptPlan>>> old;
    {
        std::lock_guard<std::mutex> lock(g_kept_mu);
        old.swap(g_kept);
    }
    // each plan is destroyed once no call holds it
    for (auto &e : old) {
        std::lock_guard<std::mutex> lock(e.second->mu);
        e.second->hp.reset();
    }
}

template <typename S>
int one_call(bool float_api, int type, int dim, int64_t M, const S *x, const S *y, const S *z,
             S *c, int isign, double eps, const int64_t *nm, int64_t N, const S *s, const S *t,
             const S *u, S *f, const wgpu_nufft_opts *opts)
{
    // counts are checked before any plan or device buffer exists
    if (M > mnufft::kMaxPoints || N > mnufft::kMaxPoints) {
        (void)(mnufft::point_count_ok(M, "point") && mnufft::point_count_ok(N, "target"));
        return WGPU_NUFFT_ERROR_PLAN;
    }
    OneCallKey key{};
    key.float_api = float_api;
    key.type = type;
    key.dim = dim;
    key.isign = isign;
    key.eps = eps;
    for (int a = 0; a < 3; ++a) key.nm[a] = nm && a < dim && dim >= 1 && dim <= 3 ? nm[a] : 0;
    std::memset(key.opts, 0, sizeof(key.opts));
    if (opts) std::memcpy(key.opts, opts, sizeof(key.opts));
    std::shared_ptr<KeptPlan> kept;
    int rc = kept_plan_for(key, nm, opts, &kept);
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(kept->mu);
    if (!kept->hp) {
        set_error("the kept plan was dropped by wgpu_nufft_shutdown");
        return WGPU_NUFFT_ERROR_INTERNAL;
    }
    HostPlan *hp = kept->hp.get();
    const S *axes[3] = {x, y, z};
    const S *taxes[3] = {s, t, u};
    rc = hp_setpts<S>(hp, M, axes, nullptr, N, taxes, nullptr);
    if (rc) return rc;
    return hp_execute<S>(hp, c, f);
}

}  // namespace

struct wgpu_nufft_plan_s {
    std::unique_ptr<HostPlan> hp;
};
struct wgpu_nufftf_plan_s {
    std::unique_ptr<HostPlan> hp;
};

extern "C" {

void wgpu_nufft_default_opts(wgpu_nufft_opts *opts)
{
    if (opts) std::memset(opts, 0, sizeof(*opts));
}

int wgpu_nufft_list_adapters(wgpu_nufft_adapter *adapters, int32_t capacity, int32_t *count)
{
    return guarded([&] {
        if (!count || (capacity > 0 && !adapters)) {
            set_error("null argument");
            return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
        }
        int n = 0;
        if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
        *count = n;
        for (int i = 0; i < n && i < capacity; ++i) {
            hipDeviceProp_t prop;
            std::memset(&adapters[i], 0, sizeof(wgpu_nufft_adapter));
            if (hipGetDeviceProperties(&prop, i) == hipSuccess) {
                std::snprintf(adapters[i].name, sizeof(adapters[i].name), "%s", prop.name);
                std::snprintf(adapters[i].pci_bus_id, sizeof(adapters[i].pci_bus_id),
                              "%04x:%02x:%02x.0", prop.pciDomainID, prop.pciBusID,
                              prop.pciDeviceID);
            }
            std::snprintf(adapters[i].backend, sizeof(adapters[i].backend), "HIP");
            std::snprintf(adapters[i].device_type, sizeof(adapters[i].device_type), "DiscreteGpu");
            adapters[i].is_default = i == 0;
        }
        return WGPU_NUFFT_SUCCESS;
    });
}

const char *wgpu_nufft_last_error(void) { return mnufft::last_error(); }

const char *wgpu_nufft_version(void) { return "0.3.2"; }

const char *wgpu_nufft_gpu_name(void)
{
    std::lock_guard<std::mutex> lock(g_name_mutex);
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        g_gpu_name.clear();
        return g_gpu_name.c_str();
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) == hipSuccess)
        g_gpu_name = std::string(prop.name) + " (HIP)";
    return g_gpu_name.c_str();
}

void wgpu_nufft_shutdown(void)
{
    try {
        drop_kept_plans();
    } catch (...) {
    }
}

int wgpu_nufft_makeplan(int32_t type, int32_t dim, const int64_t *n_modes, int32_t isign,
                        int64_t ntrans, double eps, wgpu_nufft_plan *plan,
                        const wgpu_nufft_opts *opts)
{
    return guarded([&] {
        HostPlan *hp = nullptr;
        int rc = hp_makeplan(false, type, dim, n_modes, isign, ntrans, eps, &hp, opts);
        if (plan) *plan = nullptr;
        if (rc) return rc;
        *plan = new wgpu_nufft_plan_s{std::unique_ptr<HostPlan>(hp)};
        return WGPU_NUFFT_SUCCESS;
    });
}

int wgpu_nufft_setpts(wgpu_nufft_plan plan, int64_t M, const double *x, const double *y,
                      const double *z, int64_t N, const double *s, const double *t,
                      const double *u)
{
    return guarded([&] {
        const double *axes[3] = {x, y, z}, *taxes[3] = {s, t, u};
        return hp_setpts<double>(plan ? plan->hp.get() : nullptr, M, axes, nullptr, N, taxes, nullptr);
    });
}

int wgpu_nufft_setpts_nd(wgpu_nufft_plan plan, int64_t M, const double *points, int64_t N,
                         const double *targets)
{
    return guarded([&] {
        const double *axes[3] = {points, points, points}, *taxes[3] = {targets, targets, targets};
        if (M > 0 && !points) {
            set_error("points array is null");
            return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
        }
        return hp_setpts<double>(plan ? plan->hp.get() : nullptr, M, axes, points, N, taxes, targets);
    });
}

int wgpu_nufft_execute(wgpu_nufft_plan plan, double *c, double *f)
{
    return guarded([&] { return hp_execute<double>(plan ? plan->hp.get() : nullptr, c, f); });
}

void wgpu_nufft_destroy(wgpu_nufft_plan plan)
{
    try {
        delete plan;
    } catch (...) {
    }
}

int wgpu_nufft_plan_backend(wgpu_nufft_plan plan) { return plan ? WGPU_NUFFT_BACKEND_GPU : -1; }

int wgpu_nufft_plan_precision(wgpu_nufft_plan plan)
{
    if (!plan) return -1;
    return plan->hp->precision == Precision::F32 ? WGPU_NUFFT_PRECISION_F32
                                                : WGPU_NUFFT_PRECISION_F64;
}

int wgpu_nufftf_makeplan(int32_t type, int32_t dim, const int64_t *n_modes, int32_t isign,
                         int64_t ntrans, double eps, wgpu_nufftf_plan *plan,
                         const wgpu_nufft_opts *opts)
{
    return guarded([&] {
        HostPlan *hp = nullptr;
        int rc = hp_makeplan(true, type, dim, n_modes, isign, ntrans, eps, &hp, opts);
        if (plan) *plan = nullptr;
        if (rc) return rc;
        *plan = new wgpu_nufftf_plan_s{std::unique_ptr<HostPlan>(hp)};
        return WGPU_NUFFT_SUCCESS;
    });
}

int wgpu_nufftf_setpts(wgpu_nufftf_plan plan, int64_t M, const float *x, const float *y,
                       const float *z, int64_t N, const float *s, const float *t, const float *u)
{
    return guarded([&] {
        const float *axes[3] = {x, y, z}, *taxes[3] = {s, t, u};
        return hp_setpts<float>(plan ? plan->hp.get() : nullptr, M, axes, nullptr, N, taxes, nullptr);
    });
}

int wgpu_nufftf_setpts_nd(wgpu_nufftf_plan plan, int64_t M, const float *points, int64_t N,
                          const float *targets)
{
    return guarded([&] {
        const float *axes[3] = {points, points, points}, *taxes[3] = {targets, targets, targets};
        if (M > 0 && !points) {
            set_error("points array is null");
            return WGPU_NUFFT_ERROR_INVALID_ARGUMENT;
        }
        return hp_setpts<float>(plan ? plan->hp.get() : nullptr, M, axes, points, N, taxes, targets);
    });
}

int wgpu_nufftf_execute(wgpu_nufftf_plan plan, float *c, float *f)
{
    return guarded([&] { return hp_execute<float>(plan ? plan->hp.get() : nullptr, c, f); });
}

void wgpu_nufftf_destroy(wgpu_nufftf_plan plan)
{
    try {
        delete plan;
    } catch (...) {
    }
}

int wgpu_nufftf_plan_backend(wgpu_nufftf_plan plan) { return plan ? WGPU_NUFFT_BACKEND_GPU : -1; }

int wgpu_nufftf_plan_precision(wgpu_nufftf_plan plan)
{
    return plan ? WGPU_NUFFT_PRECISION_F32 : -1;
}

int wgpu_nufft1d1(int64_t M, const double *x, const double *c, int32_t isign, double eps,
                  int64_t ms, double *f, const wgpu_nufft_opts *opts)
{
    int64_t nm[3] = {ms, 1, 1};
    return guarded([&] {
        return one_call<double>(false, 1, 1, M, x, nullptr, nullptr, const_cast<double *>(c),
                                isign, eps, nm, 0, nullptr, nullptr, nullptr, f, opts);
    });
}
int wgpu_nufft1d2(int64_t M, const double *x, double *c, int32_t isign, double eps, int64_t ms,
                  const double *f, const wgpu_nufft_opts *opts)
{
    int64_t nm[3] = {ms, 1, 1};
    return guarded([&] {
        return one_call<double>(false, 2, 1, M, x, nullptr, nullptr, c, isign, eps, nm, 0,
                                nullptr, nullptr, nullptr, const_cast<double *>(f), opts);
    });
}
int wgpu_nufft1d3(int64_t M, const double *x, const double *c, int32_t isign, double eps,
                  int64_t N, const double *s, double *f, const wgpu_nufft_opts *opts)
{
    return guarded([&] {
        return one_call<double>(false, 3, 1, M, x, nullptr, nullptr, const_cast<double *>(c),
                                isign, eps, nullptr, N, s, nullptr, nullptr, f, opts);
    });
}
int wgpu_nufft2d1(int64_t M, const double *x, const double *y, const double *c, int32_t isign,
                  double eps, int64_t ms, int64_t mt, double *f, const wgpu_nufft_opts *opts)
{
    int64_t nm[3] = {ms, mt, 1};
    return guarded([&] {
        return one_call<double>(false, 1, 2, M, x, y, nullptr, const_cast<double *>(c), isign,
                                eps, nm, 0, nullptr, nullptr, nullptr, f, opts);
    });
}
int wgpu_nufft2d2(int64_t M, const double *x, const double *y, double *c, int32_t isign,
                  double eps, int64_t ms, int64_t mt, const double *f,
                  const wgpu_nufft_opts *opts)
{
    int64_t nm[3] = {ms, mt, 1};
    return guarded([&] {
        return one_call<double>(false, 2, 2, M, x, y, nullptr, c, isign, eps, nm, 0, nullptr,
                                nullptr, nullptr, const_cast<double *>(f), opts);
    });
}
int wgpu_nufft2d3(int64_t M, const double *x, const double *y, const double *c, int32_t isign,
                  double eps, int64_t N, const double *s, const double *t, double *f,
                  const wgpu_nufft_opts *opts)
{
    return guarded([&] {
        return one_call<double>(false, 3, 2, M, x, y, nullptr, const_cast<double *>(c), isign,
                                eps, nullptr, N, s, t, nullptr, f, opts);
    });
}
int wgpu_nufft3d1(int64_t M, const double *x, const double *y, const double *z, const double *c,
                  int32_t isign, double eps, int64_t ms, int64_t mt, int64_t mu, double *f,
                  const wgpu_nufft_opts *opts)
{
    int64_t nm[3] = {ms, mt, mu};
    return guarded([&] {
        return one_call<double>(false, 1, 3, M, x, y, z, const_cast<double *>(c), isign, eps, nm,
                                0, nullptr, nullptr, nullptr, f, opts);
    });
}
int wgpu_nufft3d2(int64_t M, const double *x, const double *y, const double *z, double *c,
                  int32_t isign, double eps, int64_t ms, int64_t mt, int64_t mu,
                  const double *f, const wgpu_nufft_opts *opts)
{
    int64_t nm[3] = {ms, mt, mu};
    return guarded([&] {
        return one_call<double>(false, 2, 3, M, x, y, z, c, isign, eps, nm, 0, nullptr, nullptr,
                                nullptr, const_cast<double *>(f), opts);
    });
}
int wgpu_nufft3d3(int64_t M, const double *x, const double *y, const double *z, const double *c,
                  int32_t isign, double eps, int64_t N, const double *s, const double *t,
                  const double *u, double *f, const wgpu_nufft_opts *opts)
{
    return guarded([&] {
        return one_call<double>(false, 3, 3, M, x, y, z, const_cast<double *>(c), isign, eps,
                                nullptr, N, s, t, u, f, opts);
    });
}

int wgpu_nufftf1d1(int64_t M, const float *x, const float *c, int32_t isign, double eps,
                   int64_t ms, float *f, const wgpu_nufft_opts *opts)
{
    int64_t nm[3] = {ms, 1, 1};
    return guarded([&] {
        return one_call<float>(true, 1, 1, M, x, nullptr, nullptr, const_cast<float *>(c), isign,
                               eps, nm, 0, nullptr, nullptr, nullptr, f, opts);
    });
}
int wgpu_nufftf1d2(int64_t M, const float *x, float *c, int32_t isign, double eps, int64_t ms,
                   const float *f, const wgpu_nufft_opts *opts)
{
    int64_t nm[3] = {ms, 1, 1};
    return guarded([&] {
        return one_call<float>(true, 2, 1, M, x, nullptr, nullptr, c, isign, eps, nm, 0, nullptr,
                               nullptr, nullptr, const_cast<float *>(f), opts);
    });
}
int wgpu_nufftf1d3(int64_t M, const float *x, const float *c, int32_t isign, double eps,
                   int64_t N, const float *s, float *f, const wgpu_nufft_opts *opts)
{
    return guarded([&] {
        return one_call<float>(true, 3, 1, M, x, nullptr, nullptr, const_cast<float *>(c), isign,
                               eps, nullptr, N, s, nullptr, nullptr, f, opts);
    });
}
int wgpu_nufftf2d1(int64_t M, const float *x, const float *y, const float *c, int32_t isign,
                   double eps, int64_t ms, int64_t mt, float *f, const wgpu_nufft_opts *opts)
{
    int64_t nm[3] = {ms, mt, 1};
    return guarded([&] {
        return one_call<float>(true, 1, 2, M, x, y, nullptr, const_cast<float *>(c), isign, eps,
                               nm, 0, nullptr, nullptr, nullptr, f, opts);
    });
}
int wgpu_nufftf2d2(int64_t M, const float *x, const float *y, float *c, int32_t isign,
                   double eps, int64_t ms, int64_t mt, const float *f,
                   const wgpu_nufft_opts *opts)
{
    int64_t nm[3] = {ms, mt, 1};
    return guarded([&] {
        return one_call<float>(true, 2, 2, M, x, y, nullptr, c, isign, eps, nm, 0, nullptr,
                               nullptr, nullptr, const_cast<float *>(f), opts);
    });
}
int wgpu_nufftf2d3(int64_t M, const float *x, const float *y, const float *c, int32_t isign,
                   double eps, int64_t N, const float *s, const float *t, float *f,
                   const wgpu_nufft_opts *opts)
{
    return guarded([&] {
        return one_call<float>(true, 3, 2, M, x, y, nullptr, const_cast<float *>(c), isign, eps,
                               nullptr, N, s, t, nullptr, f, opts);
    });
}
int wgpu_nufftf3d1(int64_t M, const float *x, const float *y, const float *z, const float *c,
                   int32_t isign, double eps, int64_t ms, int64_t mt, int64_t mu, float *f,
                   const wgpu_nufft_opts *opts)
{
    int64_t nm[3] = {ms, mt, mu};
    return guarded([&] {
        return one_call<float>(true, 1, 3, M, x, y, z, const_cast<float *>(c), isign, eps, nm, 0,
                               nullptr, nullptr, nullptr, f, opts);
    });
}
int wgpu_nufftf3d2(int64_t M, const float *x, const float *y, const float *z, float *c,
                   int32_t isign, double eps, int64_t ms, int64_t mt, int64_t mu,
                   const float *f, const wgpu_nufft_opts *opts)
{
    int64_t nm[3] = {ms, mt, mu};
    return guarded([&] {
        return one_call<float>(true, 2, 3, M, x, y, z, c, isign, eps, nm, 0, nullptr, nullptr,
                               nullptr, const_cast<float *>(f), opts);
    });
}
int wgpu_nufftf3d3(int64_t M, const float *x, const float *y, const float *z, const float *c,
                   int32_t isign, double eps, int64_t N, const float *s, const float *t,
                   const float *u, float *f, const wgpu_nufft_opts *opts)
{
    return guarded([&] {
        return one_call<float>(true, 3, 3, M, x, y, z, const_cast<float *>(c), isign, eps,
                               nullptr, N, s, t, u, f, opts);
    });
}

}  // extern "C"
