// Batched gradient-based geometry calibration for gfx950.
//
// Each problem is a known volume and a measured projection stack; its eight
// per-view geometry arrays are fitted by Adam to the mean squared reprojection
// residual  L = mean over rays of (project(volume, geometry) - measured)^2.
// One optimisation step of all problems is
//   1. frames      per-view frames of every problem from the current geometry;
//   2. march       the persistent mode-3 slab kernel: every ray of every problem
//                  is marched once, giving its projection, its squared residual and
//                  the geometry gradient of the loss (no separate forward pass);
//   3. reduce      per-view sums of the per-tile partials, in a fixed order;
//   4. adam        the Adam update of every geometry value.
// The steps can be captured once in a hipGraph and launched as one graph.
// Problems of uneven size share every launch; persistent workgroups pull tiles
// from one counter, which balances the uneven work.
#include "cbct_common.h"
#include "cbct_fp_slab.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <utility>
#include <vector>

namespace cbct {

__device__ __forceinline__ int find_problem(const CalibDev* __restrict__ probs, int nprob, int idx, int which) {
    int lo = 0, hi = nprob - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int off = which == 0 ? probs[mid].view_offset : probs[mid].param_offset;
        if (off <= idx) lo = mid; else hi = mid - 1;
    }
    return lo;
}

__global__ void calib_zero_kernel(unsigned* __restrict__ counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] = 0u;
}

__global__ void calib_frames_kernel(const CalibDev* __restrict__ probs, int nprob, int total_views) {
    const int gv = blockIdx.x * blockDim.x + threadIdx.x;
    if (gv >= total_views) return;
    const int p = find_problem(probs, nprob, gv, 0);
    const CalibDev& d = probs[p];
    const int a = gv - d.view_offset, A = d.A;
    const float* g = d.geom;
    const GeomPtrs gp{g, g + A, g + 2 * A, g + 3 * A, g + 4 * A, g + 5 * A, g + 6 * A, g + 7 * A};
    // the geometry being fitted can leave the range that a march can follow (or become NaN, e.g.
    // from a NaN measurement): such a view is degenerate for this step (zero projection and
    // gradient) instead of an endless march
    const_cast<ViewFrame*>(d.s.frames)[a] = guard_frame(gp, a, d.guard);
}

// Per view: fixed-order sum of the tile partials.  NP = 9 (Adam: 8 geometry gradients + loss,
// written to ggeo / vloss and added to the loss history) or LM_NE (Levenberg-Marquardt: the
// normal equations, written to ne; the history is written by the LM kernel).
template <int NP>
__global__ __launch_bounds__(256) void calib_reduce_kernel(const CalibDev* __restrict__ probs, int nprob,
                                                           int step) {
    const int gv = blockIdx.x;
    const int p = find_problem(probs, nprob, gv, 0);
    const CalibDev& d = probs[p];
    const int a = gv - d.view_offset;
    const int tiles = d.s.tiles_u * d.s.tiles_v;
    float acc[NP];
#pragma unroll
    for (int k = 0; k < NP; ++k) acc[k] = 0.f;
    for (int t = threadIdx.x; t < tiles; t += 256) {
        const float* q = d.s.partials + ((int64_t)a * tiles + t) * NP;
#pragma unroll
        for (int k = 0; k < NP; ++k) acc[k] += q[k];
    }
    __shared__ float sm[4][NP];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < NP; ++k) {
        const float s = wave_sum(acc[k]);
        if (lane == 0) sm[wave][k] = s;
    }
    __syncthreads();
    if (threadIdx.x < NP) {
        const int k = threadIdx.x;
        const float s = (sm[0][k] + sm[1][k]) + (sm[2][k] + sm[3][k]);
        if (NP == LM_NE) {
            d.ne[(int64_t)a * LM_NE + k] = s;
            if (k == LM_NE - 1) d.vloss[a] = s;
        } else if (k < 8) {
            d.ggeo[k * d.A + a] = s;
        } else {
            d.vloss[a] = s;          // summed per problem in a fixed order by calib_loss_kernel
        }
    }
}

// Levenberg-Marquardt step per view (one thread per view).  The march of this iteration gave
// the normal equations at the current geometry; if its loss is not above the accepted one the
// current geometry is accepted (damping / 3), otherwise the view returns to the accepted
// geometry (damping x 4).  The next geometry is accepted + delta with
//   (H + lambda diag(H)) delta = -J^T r,  H = J^T J,
// solved by an FP64 Cholesky factorisation.  The loss history records the accepted loss.
__global__ void calib_lm_kernel(const CalibDev* __restrict__ probs, int nprob, int total_views, int iter,
                                float lambda0, int update) {
    const int gv = blockIdx.x * blockDim.x + threadIdx.x;
    if (gv >= total_views) return;
    const int p = find_problem(probs, nprob, gv, 0);
    const CalibDev& d = probs[p];
    const int a = gv - d.view_offset, A = d.A;
    float* st = d.lm + (int64_t)a * LM_STATE;
    const float* ne = d.ne + (int64_t)a * LM_NE;
    const float loss = ne[LM_NE - 1];
    float lambda;
    if (iter == 0 || loss <= st[8 + LM_NE + 1]) {
        for (int k = 0; k < 8; ++k) st[k] = d.geom[k * A + a];
        for (int k = 0; k < LM_NE; ++k) st[8 + k] = ne[k];
        lambda = iter == 0 ? lambda0 : fmaxf(st[8 + LM_NE] * (1.0f / 3.0f), 1e-7f);
        st[8 + LM_NE + 1] = loss;
    } else {
        lambda = fminf(st[8 + LM_NE] * 4.0f, 1e7f);
    }
    st[8 + LM_NE] = lambda;
    d.vloss[a] = st[8 + LM_NE + 1];   // the accepted loss, summed by calib_loss_kernel
    if (!update) {
        for (int k = 0; k < 8; ++k) d.geom[k * A + a] = st[k];      // finish at the accepted geometry
        return;
    }
    double M[8][8], b[8];
    double dmax = 0.0;
    {
        int idx = 0;
        for (int i = 0; i < 8; ++i)
            for (int j = i; j < 8; ++j) {
                M[i][j] = M[j][i] = (double)st[8 + idx];
                ++idx;
            }
        for (int i = 0; i < 8; ++i) {
            b[i] = -(double)st[8 + 36 + i];
            dmax = fmax(dmax, M[i][i]);
        }
    }
    const double floor_ = dmax > 0.0 ? 1e-12 * dmax : 1e-30;
    for (int i = 0; i < 8; ++i) M[i][i] += (double)lambda * fmax(M[i][i], floor_) + (M[i][i] > 0.0 ? 0.0 : floor_);
    // Cholesky M = L L^T (lower triangle in place)
    bool ok = true;
    for (int j = 0; j < 8 && ok; ++j) {
        double s = M[j][j];
        for (int k = 0; k < j; ++k) s -= M[j][k] * M[j][k];
        if (!(s > 0.0)) { ok = false; break; }
        const double l = sqrt(s);
        M[j][j] = l;
        for (int i = j + 1; i < 8; ++i) {
            double t = M[i][j];
            for (int k = 0; k < j; ++k) t -= M[i][k] * M[j][k];
            M[i][j] = t / l;
        }
    }
    double x[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    if (ok) {
        for (int i = 0; i < 8; ++i) {            // L y = b
            double t = b[i];
            for (int k = 0; k < i; ++k) t -= M[i][k] * x[k];
            x[i] = t / M[i][i];
        }
        for (int i = 7; i >= 0; --i) {           // L^T x = y
            double t = x[i];
            for (int k = i + 1; k < 8; ++k) t -= M[k][i] * x[k];
            x[i] = t / M[i][i];
        }
    }
    for (int k = 0; k < 8; ++k) d.geom[k * A + a] = (float)((double)st[k] + x[k]);
}

// loss_hist[step] of every problem: the fixed-order sum of its per-view losses (one workgroup per
// problem), so the history is bitwise reproducible and equal whatever the batch around it.
__global__ __launch_bounds__(256) void calib_loss_kernel(const CalibDev* __restrict__ probs, int step) {
    const CalibDev& d = probs[blockIdx.x];
    float acc = 0.f;
    for (int a = threadIdx.x; a < d.A; a += 256) acc += d.vloss[a];
    __shared__ float sm[4];
    acc = wave_sum(acc);
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) d.loss_hist[step] = (sm[0] + sm[1]) + (sm[2] + sm[3]);
}

// Copies up to DESC_BATCH problem descriptors, passed by value, into device memory: the upload
// needs no host buffer that outlives the call and no synchronisation.
constexpr int DESC_BATCH = 8;
struct DescBatch {
    CalibDev d[DESC_BATCH];
};
static_assert(sizeof(DescBatch) <= 3072, "descriptor batch must fit the kernel argument space");

__global__ void calib_desc_kernel(DescBatch batch, int n, CalibDev* __restrict__ out) {
    const int words = (int)(sizeof(CalibDev) / sizeof(uint32_t));
    static_assert(sizeof(CalibDev) % sizeof(uint32_t) == 0, "descriptor of whole words");
    for (int i = threadIdx.x; i < n * words; i += blockDim.x) {
        const int k = i / words, w = i % words;
        reinterpret_cast<uint32_t*>(out + k)[w] = reinterpret_cast<const uint32_t*>(&batch.d[k])[w];
    }
}

struct AdamArgs {
    float lr[8];
    float beta1, beta2, eps, bc1, bc2;   // bias corrections 1 - beta^t
};

__global__ void calib_adam_kernel(const CalibDev* __restrict__ probs, int nprob, int total_params, AdamArgs h) {
    const int gi = blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= total_params) return;
    const int p = find_problem(probs, nprob, gi, 1);
    const CalibDev& d = probs[p];
    const int li = gi - d.param_offset;
    const int k = li / d.A;
    const float g = d.ggeo[li];
    const float m = h.beta1 * d.m[li] + (1.0f - h.beta1) * g;
    const float v = h.beta2 * d.v[li] + (1.0f - h.beta2) * g * g;
    d.m[li] = m;
    d.v[li] = v;
    const float mhat = m / h.bc1, vhat = v / h.bc2;
    d.geom[li] -= h.lr[k] * mhat / (sqrtf(vhat) + h.eps);
}

static size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

struct CalibLayout {
    size_t descs, counter, total;
    std::vector<size_t> frames, partials, ggeo, vloss, ne, lm;
};

static int64_t tiles_of(int64_t U, int64_t V) { return ceil_div(U, SL_TILE) * ceil_div(V, SL_TILE); }

static CalibLayout calib_layout(const cbct_calib_problem* pr, int n, int solver) {
    CalibLayout L;
    size_t off = 0;
    const size_t np = solver == 1 ? LM_NE : 9;
    L.descs = off; off += align256(sizeof(CalibDev) * (size_t)n);
    L.counter = off; off += 256;
    for (int i = 0; i < n; ++i) {
        const int64_t A = pr[i].A;
        L.frames.push_back(off); off += align256((size_t)A * sizeof(ViewFrame));
        L.partials.push_back(off); off += align256((size_t)A * tiles_of(pr[i].U, pr[i].V) * np * sizeof(float));
        L.ggeo.push_back(off); off += align256((size_t)A * 8 * sizeof(float));
        L.vloss.push_back(off); off += align256((size_t)A * sizeof(float));
        L.ne.push_back(off); off += solver == 1 ? align256((size_t)A * LM_NE * sizeof(float)) : 0;
        L.lm.push_back(off); off += solver == 1 ? align256((size_t)A * LM_STATE * sizeof(float)) : 0;
    }
    L.total = off;
    return L;
}

// ---- instantiated-graph cache of cbct_calib_run / cbct_calib_run_lm
struct GraphKey {
    int device;
    const void* ws;
    int nproblems, nsteps, first_step, solver, total_tiles, total_views, total_params, max_cap, grid, big;
    float lr[8];
    float lr_decay, beta1, beta2, eps, lambda0;
};

constexpr size_t GRAPH_CACHE_ENTRIES = 8;
// A cached graph and an event recorded after its latest launch.
struct CachedGraph {
    GraphKey key;
    hipGraphExec_t exec;
    hipEvent_t done;
};
// Cached graphs (oldest first) and graphs evicted from the cache.  A graph is launched, and its
// event recorded, while the mutex is held (both only enqueue work); an evicted graph is destroyed
// only once its event reports that its last launch has completed (checked without waiting at later
// insertions) or by cbct_calib_release_graphs, so no graph is freed while another thread launches
// it or while it still runs.
static std::mutex g_graph_mutex;
static std::vector<CachedGraph> g_graphs;
static std::vector<CachedGraph> g_evicted;

static bool same_key(const GraphKey& a, const GraphKey& b) { return std::memcmp(&a, &b, sizeof(GraphKey)) == 0; }

static hipError_t launch_recorded(const CachedGraph& g, hipStream_t stream) {
    const hipError_t e = hipGraphLaunch(g.exec, stream);
    return e != hipSuccess ? e : hipEventRecord(g.done, stream);
}

static void destroy_cached(const CachedGraph& g, hipError_t& first) {
    hipError_t r = hipGraphExecDestroy(g.exec);
    if (first == hipSuccess) first = r;
    r = hipEventDestroy(g.done);
    if (first == hipSuccess) first = r;
}

// Launches the cached graph of key k on stream (true), or returns false when there is none.
static bool graph_cache_launch(const GraphKey& k, hipStream_t stream, hipError_t& err) {
    std::lock_guard<std::mutex> lock(g_graph_mutex);
    for (auto& g : g_graphs)
        if (same_key(g.key, k)) {
            err = launch_recorded(g, stream);
            return true;
        }
    return false;
}

// Caches a newly instantiated graph (evicting the oldest beyond GRAPH_CACHE_ENTRIES), launches it,
// and frees the evicted graphs whose last launch has completed.
static hipError_t graph_cache_insert_launch(const GraphKey& k, hipGraphExec_t exec, hipStream_t stream) {
    hipEvent_t done = nullptr;
    hipError_t e = hipEventCreateWithFlags(&done, hipEventDisableTiming);
    if (e != hipSuccess) {
        (void)hipGraphExecDestroy(exec);
        return e;
    }
    std::lock_guard<std::mutex> lock(g_graph_mutex);
    hipError_t first = hipSuccess;
    for (size_t i = 0; i < g_evicted.size();) {
        if (hipEventQuery(g_evicted[i].done) == hipSuccess) {
            destroy_cached(g_evicted[i], first);
            g_evicted.erase(g_evicted.begin() + (long)i);
        } else {
            ++i;
        }
    }
    if (g_graphs.size() >= GRAPH_CACHE_ENTRIES) {
        g_evicted.push_back(g_graphs.front());
        g_graphs.erase(g_graphs.begin());
    }
    g_graphs.push_back(CachedGraph{k, exec, done});
    e = launch_recorded(g_graphs.back(), stream);
    return e != hipSuccess ? e : first;
}

static int graph_cache_clear() {
    std::lock_guard<std::mutex> lock(g_graph_mutex);
    hipError_t first = hipSuccess;
    for (auto& g : g_graphs) destroy_cached(g, first);
    for (auto& g : g_evicted) destroy_cached(g, first);
    g_graphs.clear();
    g_evicted.clear();
    if (first != hipSuccess) return hip_status(first, "graph release");
    return CBCT_OK;
}

}  // namespace cbct

using namespace cbct;

extern "C" size_t cbct_calib_workspace_bytes(const cbct_calib_problem* problems, int nproblems) {
    if (!problems || nproblems < 1) return 0;
    return calib_layout(problems, nproblems, 0).total;
}

extern "C" size_t cbct_calib_lm_workspace_bytes(const cbct_calib_problem* problems, int nproblems) {
    if (!problems || nproblems < 1) return 0;
    return calib_layout(problems, nproblems, 1).total;
}

static int run_calibration(const cbct_calib_problem* pr, int nproblems, int nsteps, int first_step, int solver,
                           const float* lr8, float lr_decay, float beta1, float beta2, float eps, float lambda0,
                           int use_graph, void* workspace, size_t workspace_bytes, void* stream_);

extern "C" int cbct_calib_run(const cbct_calib_problem* pr, int nproblems, int nsteps, int first_step,
                              const float* lr8, float lr_decay, float beta1, float beta2, float eps, int use_graph,
                              void* workspace, size_t workspace_bytes, void* stream_) {
    CBCT_CHECK_ARG(pr != nullptr && nproblems >= 1, "at least one problem is required");
    CBCT_CHECK_ARG(nsteps >= 0 && nsteps <= 100000, "nsteps must be in [0, 100000]");
    CBCT_CHECK_ARG(first_step >= 0 && first_step <= 100000000, "first_step must be in [0, 1e8]");
    CBCT_CHECK_ARG(lr_decay > 0.f && lr_decay <= 1.f, "lr_decay must be in (0, 1]");
    CBCT_CHECK_ARG(lr8 != nullptr, "lr must point to 8 learning rates");
    CBCT_CHECK_ARG(beta1 >= 0.f && beta1 < 1.f && beta2 >= 0.f && beta2 < 1.f && eps > 0.f,
                   "Adam parameters out of range");
    return run_calibration(pr, nproblems, nsteps, first_step, 0, lr8, lr_decay, beta1, beta2, eps, 0.f, use_graph,
                           workspace, workspace_bytes, stream_);
}

extern "C" int cbct_calib_run_lm(const cbct_calib_problem* pr, int nproblems, int nsteps, float lambda0,
                                 int use_graph, void* workspace, size_t workspace_bytes, void* stream_) {
    CBCT_CHECK_ARG(pr != nullptr && nproblems >= 1, "at least one problem is required");
    CBCT_CHECK_ARG(nsteps >= 0 && nsteps <= 100000, "nsteps must be in [0, 100000]");
    CBCT_CHECK_ARG(lambda0 > 0.f && lambda0 <= 1e6f, "lambda0 must be in (0, 1e6]");
    return run_calibration(pr, nproblems, nsteps, 0, 1, nullptr, 1.f, 0.f, 0.f, 1.f, lambda0, use_graph,
                           workspace, workspace_bytes, stream_);
}

static int run_calibration(const cbct_calib_problem* pr, int nproblems, int nsteps, int first_step, int solver,
                           const float* lr8, float lr_decay, float beta1, float beta2, float eps, float lambda0,
                           int use_graph, void* workspace, size_t workspace_bytes, void* stream_) {
    hipStream_t stream = (hipStream_t)stream_;
    const CalibLayout L = calib_layout(pr, nproblems, solver);
    CBCT_CHECK_ARG(workspace != nullptr && workspace_bytes >= L.total, "workspace too small for cbct_calib_run");
    char* ws = (char*)workspace;
    int dflt[4] = {1, 8, 4096, 1};
    cbct_fp_get_tuning(dflt);
    std::vector<CalibDev> host(nproblems);
    int total_tiles = 0, total_views = 0, total_params = 0, max_cap = 256;
    bool big = false;
    for (int i = 0; i < nproblems; ++i) {
        const cbct_calib_problem& q = pr[i];
        CBCT_CHECK_ARG(q.volume && q.measured && q.geometry && q.loss_history && (solver == 1 || (q.adam_m && q.adam_v)),
                       "problem pointers must be non-null");
        CBCT_CHECK_ARG(q.D > 0 && q.H > 0 && q.W > 0 && q.D < (1 << 24) && q.H < (1 << 24) && q.W < (1 << 24) &&
                           q.H * q.W <= INT32_MAX && q.D * q.H <= INT32_MAX,
                       "volume dimensions must be positive, below 2^24 and with planes within 2^31-1 elements");
        CBCT_CHECK_ARG(q.A > 0 && q.U > 0 && q.V > 0 && q.A * tiles_of(q.U, q.V) <= INT32_MAX,
                       "projection dimensions must be positive");
        int st;
        if ((st = check_positive_finite(q.pixel_size, "pixel_size"))) return st;
        if ((st = check_positive_finite(q.step_size, "step_size"))) return st;
        if ((st = check_positive_finite(q.voxel_size, "voxel_size"))) return st;
        if ((st = check_finite(q.off_x, "off_x")) || (st = check_finite(q.off_y, "off_y")) ||
            (st = check_finite(q.off_z, "off_z")))
            return st;
        const FrameGuard guard = make_guard_host(q.D, q.H, q.W, q.voxel_size, q.off_x, q.off_y, q.off_z, q.U, q.V,
                                                 q.pixel_size, q.step_size);
        if ((st = check_step_resolution(q.step_size, guard))) return st;
        CalibDev& d = host[i];
        SlabArgs& s = d.s;
        s.vol = q.volume;
        s.grad = q.measured;
        s.frames = (const ViewFrame*)(ws + L.frames[i]);
        s.proj = nullptr;
        s.gvol = nullptr;
        s.partials = (float*)(ws + L.partials[i]);
        s.b = make_box_host(q.D, q.H, q.W, q.voxel_size, q.off_x, q.off_y, q.off_z);
        s.U = (int)q.U; s.V = (int)q.V;
        s.tiles_u = (int)ceil_div(q.U, SL_TILE); s.tiles_v = (int)ceil_div(q.V, SL_TILE);
        s.nblocks = (int)(q.A * s.tiles_u * s.tiles_v);
        s.pix = q.pixel_size; s.step = q.step_size; s.vs = q.voxel_size; s.inv_vs = 1.0f / q.voxel_size;
        s.T = q.slab_thickness > 0 ? q.slab_thickness : dflt[1];
        s.cap = q.lds_floats > 0 ? q.lds_floats : dflt[2];
        s.layout = q.layout >= 0 && q.layout <= 1 && q.slab_thickness > 0 ? q.layout : dflt[3];
        CBCT_CHECK_ARG(s.T >= 1 && s.T <= 64 && s.cap >= 256 && s.cap <= 8192, "slab tuning out of range");
        s.resid_scale = (float)(2.0 / ((double)q.A * (double)q.U * (double)q.V));
        d.geom = q.geometry; d.m = q.adam_m; d.v = q.adam_v;
        d.ggeo = q.grad_out ? q.grad_out : (float*)(ws + L.ggeo[i]);
        d.vloss = (float*)(ws + L.vloss[i]);
        d.loss_hist = q.loss_history;
        d.A = (int)q.A;
        d.div = is_pow2(q.voxel_size) ? 0 : 1;
        d.solver = solver;
        d.ne = solver == 1 ? (float*)(ws + L.ne[i]) : nullptr;
        d.lm = solver == 1 ? (float*)(ws + L.lm[i]) : nullptr;
        d.guard = guard;
        d.view_offset = total_views;
        d.tile_offset = total_tiles;
        d.param_offset = total_params;
        total_views += (int)q.A;
        total_tiles += s.nblocks;
        total_params += 8 * (int)q.A;
        max_cap = std::max(max_cap, s.cap);
        big = big || slab_wide_offsets(q.D, q.H, q.W);
    }
    CalibDev* ddev = (CalibDev*)(ws + L.descs);
    unsigned* counter = (unsigned*)(ws + L.counter);
    int device = 0, st0;
    if ((st0 = current_stream_device(stream, &device))) return st0;   // graphs are captured on it
    // descriptors: written by kernels whose arguments carry them, ordered on the stream before
    // every step (no host synchronisation)
    for (int i0 = 0; i0 < nproblems; i0 += DESC_BATCH) {
        DescBatch batch;
        const int n = std::min(DESC_BATCH, nproblems - i0);
        for (int k = 0; k < n; ++k) batch.d[k] = host[i0 + k];
        calib_desc_kernel<<<1, 256, 0, stream>>>(batch, n, ddev + i0);
        CBCT_LAUNCHED("calibration descriptor upload");
    }
    int cus = 0;
    CBCT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device), "CU count");
    const int grid = std::max(1, std::min(total_tiles, cus * 8));
    AdamArgs h;
    for (int k = 0; k < 8; ++k) h.lr[k] = lr8 ? lr8[k] : 0.f;
    h.beta1 = beta1; h.beta2 = beta2; h.eps = eps;

    auto enqueue = [&](int step, bool update, hipStream_t st_) -> int {
        calib_zero_kernel<<<1, 64, 0, st_>>>(counter);
        CBCT_LAUNCHED("calibration counter reset");
        calib_frames_kernel<<<(unsigned)ceil_div(total_views, 128), 128, 0, st_>>>(ddev, nproblems, total_views);
        CBCT_LAUNCHED("calibration frames kernel launch");
        int st;
        if ((st = launch_fp_slab_batched(ddev, nproblems, total_tiles, max_cap, solver, counter, grid, st_, big))) return st;
        if (solver == 1) {
            calib_reduce_kernel<LM_NE><<<(unsigned)total_views, 256, 0, st_>>>(ddev, nproblems, step);
            CBCT_LAUNCHED("calibration reduce kernel launch");
            calib_lm_kernel<<<(unsigned)ceil_div(total_views, 64), 64, 0, st_>>>(ddev, nproblems, total_views, step,
                                                                                 lambda0, update ? 1 : 0);
            CBCT_LAUNCHED("calibration LM kernel launch");
            calib_loss_kernel<<<(unsigned)nproblems, 256, 0, st_>>>(ddev, step);
            CBCT_LAUNCHED("calibration loss kernel launch");
            return CBCT_OK;
        }
        calib_reduce_kernel<9><<<(unsigned)total_views, 256, 0, st_>>>(ddev, nproblems, step);
        CBCT_LAUNCHED("calibration reduce kernel launch");
        calib_loss_kernel<<<(unsigned)nproblems, 256, 0, st_>>>(ddev, step);
        CBCT_LAUNCHED("calibration loss kernel launch");
        if (update) {
            const int t = first_step + step;
            const double decay = std::pow((double)lr_decay, t);
            for (int k = 0; k < 8; ++k) h.lr[k] = (float)(lr8[k] * decay);
            h.bc1 = (float)(1.0 - std::pow((double)beta1, t + 1));
            h.bc2 = (float)(1.0 - std::pow((double)beta2, t + 1));
            calib_adam_kernel<<<(unsigned)ceil_div(total_params, 256), 256, 0, st_>>>(ddev, nproblems,
                                                                                      total_params, h);
            CBCT_LAUNCHED("calibration Adam kernel launch");
        }
        return CBCT_OK;
    };

    if (!use_graph) {
        for (int s = 0; s < nsteps; ++s) {
            int st;
            if ((st = enqueue(s, true, stream))) return st;
        }
        return enqueue(nsteps, false, stream);
    }
    // All steps in one graph.  Its kernels read every per-problem pointer from the descriptors in
    // the workspace (rewritten above on every call), so an instantiated graph depends only on the
    // key below and is kept for later calls with the same key (cbct_calib_release_graphs frees them).
    GraphKey key;
    std::memset(&key, 0, sizeof(key));    // padding included, the key is compared bytewise
    key.device = device; key.ws = ws; key.nproblems = nproblems; key.nsteps = nsteps; key.first_step = first_step;
    key.solver = solver; key.total_tiles = total_tiles; key.total_views = total_views;
    key.total_params = total_params; key.max_cap = max_cap; key.grid = grid; key.big = big ? 1 : 0;
    for (int k = 0; k < 8; ++k) key.lr[k] = lr8 ? lr8[k] : 0.f;
    key.lr_decay = lr_decay; key.beta1 = beta1; key.beta2 = beta2; key.eps = eps; key.lambda0 = lambda0;
    hipError_t launched = hipSuccess;
    if (graph_cache_launch(key, stream, launched)) {
        CBCT_HIP(launched, "graph launch");
        return CBCT_OK;
    }
    // capture on a private stream (the caller's may be the null stream, which cannot be captured)
    hipStream_t cs = nullptr;
    CBCT_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "capture stream");
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) { (void)hipStreamDestroy(cs); return hip_status(e, "graph capture begin"); }
    int st = CBCT_OK;
    for (int s = 0; s < nsteps && st == CBCT_OK; ++s) st = enqueue(s, true, cs);
    if (st == CBCT_OK) st = enqueue(nsteps, false, cs);
    e = hipStreamEndCapture(cs, &graph);
    (void)hipStreamDestroy(cs);
    if (st != CBCT_OK) { if (graph) (void)hipGraphDestroy(graph); return st; }
    if (e != hipSuccess) return hip_status(e, "graph capture end");
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return hip_status(e, "graph instantiate");
    CBCT_HIP(graph_cache_insert_launch(key, exec, stream), "graph launch");
    return CBCT_OK;
}

extern "C" int cbct_calib_release_graphs(void) {
    return graph_cache_clear();
}
