// hfm_scene.hip — C ABI of libhfm.so and the host-side order of the substep.
//
// HeatflowMPM.substep (solver.py:128-151) as launched here, one HIP stream per scene:
//   [particles] bin (when positions changed) + deterministic P2G gather
//   heat_prep -> advect_T -> heat PCG (x = T2 in place) -> [particles] G2P -> swap T, T2
//   air_faces -> divergence -> numpy-exact rhs mean -> pressure PCG (warm start) -> project
// The host never waits for the GPU inside a substep except to decide how many PCG iterations
// to enqueue: the last workgroup of every iteration publishes its progress to pinned host
// memory, and the host keeps a bounded number of iterations in flight.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <string>
#include <thread>
#include <vector>

#include "hfm_device.h"
#include "hfm_scene.h"

using namespace hfm;

namespace {

thread_local std::string g_err;

void set_err(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

#define HIP_OK(expr)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess) {                                                                  \
            set_err("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, __LINE__); \
            return e_ == hipErrorOutOfMemory ? HFM_ERR_OOM : HFM_ERR_HIP;                        \
        }                                                                                        \
    } while (0)

#define CHECK_LAUNCH()                                                                       \
    do {                                                                                     \
        hipError_t e_ = hipGetLastError();                                                   \
        if (e_ != hipSuccess) {                                                              \
            set_err("kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), __FILE__, __LINE__); \
            return HFM_ERR_HIP;                                                              \
        }                                                                                    \
    } while (0)

struct Owner {
    Scene s;
    std::vector<void*> dev;
    std::vector<int> prog;
};

inline Owner* owner(hfm_scene* h) { return reinterpret_cast<Owner*>(h); }

template <class T>
int dalloc(Owner* o, T** p, long count) {
    size_t bytes = (size_t)(count > 0 ? count : 1) * sizeof(T);
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) {
        set_err("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? HFM_ERR_OOM : HFM_ERR_HIP;
    }
    o->dev.push_back(q);
    HIP_OK(hipMemset(q, 0, bytes));
    *p = (T*)q;
    return HFM_OK;
}

void dfree(Owner* o, void* p) {
    for (auto& q : o->dev)
        if (q == p) {
            (void)hipFree(q);
            q = nullptr;
        }
}

// numpy pairwise-sum tree of one buffer chunk: leaves of at most 128 elements, combined by
// splitting at n/2 rounded down to a multiple of 8 (numpy loops_utils pairwise_sum).
void pairwise_program(int off, int len, std::vector<int>& leaves, std::vector<int>& ops) {
    if (len <= 128) {
        ops.push_back((int)(leaves.size() / 2));
        leaves.push_back(off);
        leaves.push_back(len);
        return;
    }
    int n2 = len / 2;
    n2 -= n2 % 8;
    pairwise_program(off, n2, leaves, ops);
    pairwise_program(off + n2, len - n2, leaves, ops);
    ops.push_back(-1);
}

void append_program(std::vector<int>& out, int len) {
    std::vector<int> leaves, ops;
    pairwise_program(0, len, leaves, ops);
    out.push_back((int)(leaves.size() / 2));
    out.push_back((int)ops.size());
    out.insert(out.end(), leaves.begin(), leaves.end());
    out.insert(out.end(), ops.begin(), ops.end());
}

int warp_reported_count(int exact, int converged, int maxiter, int check_every) {
    int cur = 0;
    for (;;) {
        cur += 2;
        if (cur >= maxiter) break;
        if ((cur % check_every) < 2 && converged && exact <= cur) break;
    }
    return cur;
}

// ---- kernel timing -------------------------------------------------------------------------

int flush_timing(Scene* s) {
    if (!s->timing || s->ev_used == 0) return HFM_OK;
    HIP_OK(hipStreamSynchronize(s->stream));
    for (int e = 0; e + 1 < s->ev_used; e += 2) {
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, s->ev_pool[e], s->ev_pool[e + 1]));
        s->kt_ms[s->ev_class[e]] += ms;
        s->kt_n[s->ev_class[e]] += 1;
    }
    s->ev_used = 0;
    return HFM_OK;
}

// A pair of events for one timed launch of kernel class cls (flushing the pool when full).
int timing_events(Scene* s, int cls, hipEvent_t* ev0, hipEvent_t* ev1) {
    *ev0 = *ev1 = nullptr;
    if (!s->timing) return HFM_OK;
    if (s->ev_used + 2 > s->ev_cap) {
        int rc = flush_timing(s);
        if (rc) return rc;
    }
    s->ev_class[s->ev_used] = cls;
    *ev0 = s->ev_pool[s->ev_used++];
    *ev1 = s->ev_pool[s->ev_used++];
    return HFM_OK;
}

int enable_timing(Scene* s) {
    if (!s->ev_pool) {
        s->ev_cap = 4096;
        s->ev_pool = new hipEvent_t[s->ev_cap];
        s->ev_class = new int[s->ev_cap];
        for (int i = 0; i < s->ev_cap; ++i) HIP_OK(hipEventCreate(&s->ev_pool[i]));
    }
    s->timing = 1;
    return HFM_OK;
}

// ---- PCG driver ----------------------------------------------------------------------------

// Enqueue one PCG solve (op 0 heat, 1 pressure) and return when the device reports it done.
int run_pcg(Scene* s, int op, float dt) {
    CgMirror* m = &s->mirror[op];
    __atomic_store_n(&m->it, 0, __ATOMIC_RELAXED);
    __atomic_store_n(&m->done, 0, __ATOMIC_RELEASE);
    launch_pcg_init(s, op, dt);
    CHECK_LAUNCH();
    const int maxiter = op ? s->prm.cg_max_iter : s->prm.heat_max_iter;
    const int maxeff = maxiter + (maxiter & 1);
    int& pred = op ? s->pres_pred : s->heat_pred;
    // Iterations are enqueued in batches of (pass A, pass B) pairs, at most two batches in
    // flight.  Pass A of iteration k+1 finishes iteration k, so the first pass A of a batch
    // posts how many iterations the previous batch completed, which releases the next batch;
    // after the last pair one more pass A closes the solve.  Batch size follows the iteration
    // count of the previous solve so that few launches are wasted after convergence.
    int launched = 0;
    int batch_end[2] = {0, 0};
    int inflight = 0;
    bool closed = false;
    auto t0 = std::chrono::steady_clock::now();
    long spins = 0;
    for (;;) {
        if (__atomic_load_n(&m->done, __ATOMIC_ACQUIRE)) break;
        int prog = __atomic_load_n(&m->it, __ATOMIC_ACQUIRE);
        while (inflight > 0 && prog >= batch_end[0]) {
            batch_end[0] = batch_end[1];
            --inflight;
        }
        if (inflight < 2 && launched < maxeff) {
            int g = pred - launched;
            g = g < 2 ? 2 : (g > 16 ? 16 : g);
            if (g > maxeff - launched) g = maxeff - launched;
            const bool timed = op == 1 && s->timing;
            for (int q = 0; q < g; ++q) {
                int k = launched + q;
                hipEvent_t e0 = nullptr, e1 = nullptr;
                int rc = timed ? timing_events(s, HFM_KT_PCG_DIR, &e0, &e1) : HFM_OK;
                if (rc) return rc;
                launch_pcg_dir(s, op, dt, k, q == 0, e0, e1);
                rc = timed ? timing_events(s, HFM_KT_PCG_UPD, &e0, &e1) : HFM_OK;
                if (rc) return rc;
                launch_pcg_upd(s, op, k, e0, e1);
            }
            CHECK_LAUNCH();
            launched += g;
            batch_end[inflight++] = launched;
            continue;
        }
        if (launched >= maxeff && !closed) {
            launch_pcg_dir(s, op, dt, maxeff, 1);
            CHECK_LAUNCH();
            closed = true;
            continue;
        }
        if ((++spins & 0x3ff) == 0) {
            double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (el > 30.0) {
                set_err("PCG (%s) did not finish within 30 s (iteration %d of %d launched)",
                        op ? "pressure" : "heat", prog, launched);
                return HFM_ERR_HIP;
            }
            hipError_t e = hipStreamQuery(s->stream);
            if (e != hipSuccess && e != hipErrorNotReady) {
                set_err("stream error during PCG: %s", hipGetErrorString(e));
                return HFM_ERR_HIP;
            }
            std::this_thread::yield();
        }
    }
    int it = __atomic_load_n(&m->it, __ATOMIC_ACQUIRE);
    int conv = __atomic_load_n(&m->converged, __ATOMIC_ACQUIRE);
    pred = it > 2 ? it : 2;
    int ce = op ? s->prm.cg_check_every : s->prm.heat_check_every;
    int reported = it == 0 && conv ? 0 : warp_reported_count(it, conv, maxiter, ce);
    if (op == 1) {
        s->stats.p_iters_total += it;
        s->stats.p_iters_exact = it;
        s->stats.p_iters = reported;
        s->stats.p_converged = conv;
        s->stats.p_rr = m->rr;
        s->stats.p_atol = m->atol;
        s->stats.rhs_mean = m->mean;
    } else {
        s->stats.h_iters_total += it;
        s->stats.h_iters_exact = it;
        s->stats.h_iters = reported;
        s->stats.h_converged = conv;
        s->stats.h_rr = m->rr;
        s->stats.h_atol = m->atol;
    }
    return HFM_OK;
}

int substep(Owner* o, double dt) {
    Scene* s = &o->s;
    const float dt32 = (float)dt;
    if (s->P > 0) {
        if (s->bins_dirty) launch_bin_particles(s);
        launch_p2g(s);
    }
    launch_heat_prep(s, dt32);
    launch_advect_T(s, dt32);
    CHECK_LAUNCH();
    int rc = run_pcg(s, 0, dt32);
    if (rc) return rc;
    if (s->P > 0) launch_g2p(s);
    float* tmp = s->T;
    s->T = s->T2;
    s->T2 = tmp;
    launch_air_faces(s, dt32);
    launch_divergence(s, (const float* const*)s->us, (float)(s->prm.rho / dt), s->rhs, 1);
    const int* pf = o->s.mean_prog;
    const int* pt = s->mean_prog_tail >= 0 ? s->mean_prog + s->mean_prog_tail : nullptr;
    launch_rhs_mean(s, pf, pt);
    CHECK_LAUNCH();
    rc = run_pcg(s, 1, dt32);
    if (rc) return rc;
    launch_project(s, (float)(dt / s->prm.rho));
    CHECK_LAUNCH();
    for (int a = 0; a < 3; ++a) {
        float* t = s->u[a];
        s->u[a] = s->us[a];
        s->us[a] = t;
    }
    s->stats.substeps += 1;
    return HFM_OK;
}

float* field_ptr(Scene* s, int f, long* size) {
    long N = s->g.N, F = s->g.F;
    switch (f) {
        case HFM_FIELD_T: *size = N; return s->T;
        case HFM_FIELD_T2: *size = N; return s->T2;
        case HFM_FIELD_TW: *size = N; return s->Tw;
        case HFM_FIELD_U: *size = F; return s->u[0];
        case HFM_FIELD_V: *size = F; return s->u[1];
        case HFM_FIELD_W: *size = F; return s->u[2];
        case HFM_FIELD_P: *size = N; return s->p;
        case HFM_FIELD_RHS: *size = N; return s->rhs;
        case HFM_FIELD_ALPHA_S: *size = N; return s->alpha_s;
        case HFM_FIELD_ALPHA_P: *size = N; return s->alpha_p;
        case HFM_FIELD_TSUM: *size = N; return s->tsum;
        case HFM_FIELD_KAPPA: *size = N; return s->kap;
        case HFM_FIELD_QX: case HFM_FIELD_QY: case HFM_FIELD_QZ: *size = F; return s->q[f - HFM_FIELD_QX];
        case HFM_FIELD_KX: case HFM_FIELD_KY: case HFM_FIELD_KZ: *size = F; return s->kf[f - HFM_FIELD_KX];
        case HFM_FIELD_MX: case HFM_FIELD_MY: case HFM_FIELD_MZ: *size = F; return s->mf[f - HFM_FIELD_MX];
        case HFM_FIELD_PDIAG: *size = N; return s->pdiag;
        case HFM_FIELD_HRHS: *size = N; return s->hrhs;
        case HFM_FIELD_HDIAG: *size = N; return s->hdiag;
        default: *size = 0; return nullptr;
    }
}

int check_finite(const float* a, long n, const char* what) {
    for (long i = 0; i < n; ++i)
        if (!(a[i] == a[i]) || a[i] == __builtin_inff() || a[i] == -__builtin_inff()) {
            set_err("%s[%ld] is not finite", what, i);
            return HFM_ERR_NONFINITE;
        }
    return HFM_OK;
}

// totals without particles: solid fraction and face weights are the static ones
int reset_particle_fields(Scene* s) {
    long N = s->g.N, F = s->g.F;
    HIP_OK(hipMemsetAsync(s->alpha_p, 0, N * sizeof(float), s->stream));
    HIP_OK(hipMemsetAsync(s->tsum, 0, N * sizeof(float), s->stream));
    HIP_OK(hipMemcpyAsync(s->alpha_s, s->st_alpha, N * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
    for (int a = 0; a < 3; ++a) {
        HIP_OK(hipMemsetAsync(s->q[a], 0, F * sizeof(float), s->stream));
        HIP_OK(hipMemcpyAsync(s->kf[a], s->st_k[a], F * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
        HIP_OK(hipMemcpyAsync(s->mf[a], s->st_m[a], F * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
    }
    return HFM_OK;
}

#define SCENE_OR_FAIL(h)                          \
    if (!(h)) {                                   \
        set_err("null scene");                    \
        return HFM_ERR_ARG;                       \
    }                                             \
    Owner* o = owner(h);                          \
    Scene* s = &o->s;                             \
    HIP_OK(hipSetDevice(s->device));

}  // namespace

extern "C" {

const char* hfm_last_error(void) { return g_err.c_str(); }

int hfm_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int hfm_scene_create(const hfm_params* prm, int device, hfm_scene** out) {
    if (!prm || !out) {
        set_err("null argument");
        return HFM_ERR_ARG;
    }
    *out = nullptr;
    if (prm->n < 2 || prm->n > 1024 || !(prm->h > 0) || prm->cg_max_iter < 0 || prm->heat_max_iter < 0 ||
        prm->cg_check_every < 1 || prm->heat_check_every < 1) {
        set_err("bad parameters (n=%d h=%g)", prm->n, prm->h);
        return HFM_ERR_ARG;
    }
    if (prm->precision != HFM_PREC_FP32) {
        set_err("precision mode %d is not available", prm->precision);
        return HFM_ERR_ARG;
    }
    int ndev = hfm_device_count();
    if (device < 0 || device >= ndev) {
        set_err("device %d not available (%d HIP devices)", device, ndev);
        return HFM_ERR_ARG;
    }
    HIP_OK(hipSetDevice(device));
    Owner* o = new Owner();
    Scene* s = &o->s;
    memset(s, 0, sizeof(Scene));
    s->prm = *prm;
    s->device = device;
    s->g.n = prm->n;
    s->g.N = (long)prm->n * prm->n * prm->n;
    s->g.F = (long)(prm->n + 1) * prm->n * prm->n;
    s->g.h = (float)prm->h;
    s->heat_pred = 2;
    {
        // minimum tiles per PCG workgroup (the workgroup cap may raise it); HFM_PCG_CPT overrides
        const char* e = getenv("HFM_PCG_CPT");
        int c = e ? atoi(e) : 1;
        s->pcg_cpt = c >= 1 && c <= 64 ? c : 1;
        e = getenv("HFM_PCG_FLAT_DIR");
        s->flat_dir = e && atoi(e) != 0;
        e = getenv("HFM_DIR_ICHUNK");
        int ic = e ? atoi(e) : 0;
        s->dir_ichunk = ic >= 1 && ic <= 1024 ? ic : 0;
    }
    s->pres_pred = 8;
    s->mean_prog_tail = -1;
    int rc = HFM_OK;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        set_err("hipStreamCreate: %s", hipGetErrorString(e));
        delete o;
        return HFM_ERR_HIP;
    }
    const long N = s->g.N, F = s->g.F;
#define A(ptr, cnt)                        \
    if ((rc = dalloc(o, &(ptr), (cnt)))) { \
        hfm_scene_destroy((hfm_scene*)o);  \
        return rc;                         \
    }
    for (int a = 0; a < 3; ++a) {
        A(s->u[a], F);
        A(s->us[a], F);
        A(s->q[a], F);
        A(s->kf[a], F);
        A(s->mf[a], F);
        A(s->st_k[a], F);
        A(s->st_m[a], F);
    }
    A(s->p, N);
    A(s->rhs, N);
    A(s->pdiag, N);
    A(s->alpha_s, N);
    A(s->st_alpha, N);
    A(s->T, N);
    A(s->T2, N);
    A(s->Tw, N);
    A(s->kap, N);
    A(s->alpha_p, N);
    A(s->tsum, N);
    A(s->tfix, N);
    A(s->hrhs, N);
    A(s->hdiag, N);
    A(s->flags, N);
    A(s->cg_r, N);
    A(s->cg_z, N);
    A(s->cg_d, 2 * N);
    A(s->cg_Ap, N);
    s->max_blocks = kMaxPcgBlocks;
    A(s->partials, kMaxDirBlocks + 3 * kMaxPcgBlocks);
    A(s->ctl, 2);
    long nchunks = (N + 8191) / 8192;
    A(s->chunk_sums, nchunks);
    A(s->scratch_field, F);
    A(s->umax, 4);
    A(s->bin_cnt, N + 1);
    A(s->bin_start, N + 1);
    A(s->scan_tmp, (N + 1) / 2048 + 16);
    // pairwise-sum programs for the numpy mean
    append_program(o->prog, (int)(N >= 8192 ? 8192 : N));
    if (N % 8192 != 0 && N > 8192) {
        s->mean_prog_tail = (int)o->prog.size();
        append_program(o->prog, (int)(N % 8192));
    }
    A(s->mean_prog, (long)o->prog.size());
#undef A
    e = hipMemcpy(s->mean_prog, o->prog.data(), o->prog.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_err("hipMemcpy: %s", hipGetErrorString(e));
        hfm_scene_destroy((hfm_scene*)o);
        return HFM_ERR_HIP;
    }
    // CG control blocks: tolerances and caps
    CgCtl ctl[2];
    memset(ctl, 0, sizeof(ctl));
    ctl[0].tol = prm->heat_tol;
    ctl[0].maxeff = prm->heat_max_iter + (prm->heat_max_iter & 1);
    ctl[1].tol = prm->cg_tol;
    ctl[1].maxeff = prm->cg_max_iter + (prm->cg_max_iter & 1);
    e = hipMemcpy(s->ctl, ctl, sizeof(ctl), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&s->mirror, 2 * sizeof(CgMirror), hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) {
        memset(s->mirror, 0, 2 * sizeof(CgMirror));
        e = hipHostGetDevicePointer((void**)&s->mirror_dev, s->mirror, 0);
    }
    if (e != hipSuccess) {
        set_err("control block setup: %s", hipGetErrorString(e));
        hfm_scene_destroy((hfm_scene*)o);
        return HFM_ERR_HIP;
    }
    // heat.py:182-183: T, T2, Tw start at the room temperature
    std::vector<float> tamb(N, (float)prm->t_amb);
    e = hipMemcpy(s->T, tamb.data(), N * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->T2, tamb.data(), N * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->Tw, tamb.data(), N * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_err("initial temperature: %s", hipGetErrorString(e));
        hfm_scene_destroy((hfm_scene*)o);
        return HFM_ERR_HIP;
    }
    if (prm->flags & HFM_FLAG_TIME_KERNELS) {
        int rc2 = enable_timing(s);
        if (rc2) {
            hfm_scene_destroy((hfm_scene*)o);
            return rc2;
        }
    }
    *out = (hfm_scene*)o;
    return HFM_OK;
}

int hfm_scene_destroy(hfm_scene* h) {
    if (!h) return HFM_OK;
    Owner* o = owner(h);
    Scene* s = &o->s;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    for (void* p : o->dev)
        if (p) (void)hipFree(p);
    if (s->mirror) (void)hipHostFree(s->mirror);
    if (s->ev_pool) {
        for (int i = 0; i < s->ev_cap; ++i) (void)hipEventDestroy(s->ev_pool[i]);
        delete[] s->ev_pool;
        delete[] s->ev_class;
    }
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete o;
    return HFM_OK;
}

int hfm_set_static(hfm_scene* h, const float* alpha, const float* kx, const float* ky, const float* kz,
                   const float* mx, const float* my, const float* mz) {
    SCENE_OR_FAIL(h);
    if (!alpha || !kx || !ky || !kz || !mx || !my || !mz) {
        set_err("null array");
        return HFM_ERR_ARG;
    }
    const long N = s->g.N, F = s->g.F;
    const float* k[3] = {kx, ky, kz};
    const float* m[3] = {mx, my, mz};
    int rc = check_finite(alpha, N, "alpha");
    for (int a = 0; a < 3 && !rc; ++a) {
        rc = check_finite(k[a], F, "k");
        if (!rc) rc = check_finite(m[a], F, "m");
    }
    if (rc) return rc;
    HIP_OK(hipStreamSynchronize(s->stream));
    HIP_OK(hipMemcpy(s->st_alpha, alpha, N * sizeof(float), hipMemcpyHostToDevice));
    for (int a = 0; a < 3; ++a) {
        HIP_OK(hipMemcpy(s->st_k[a], k[a], F * sizeof(float), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->st_m[a], m[a], F * sizeof(float), hipMemcpyHostToDevice));
    }
    s->has_static = 1;
    if (s->P == 0) {
        rc = reset_particle_fields(s);
        if (rc) return rc;
        HIP_OK(hipStreamSynchronize(s->stream));
    }
    return HFM_OK;
}

int hfm_set_fixed(hfm_scene* h, const int32_t* fixed, const float* tfix) {
    SCENE_OR_FAIL(h);
    if (!fixed || !tfix) {
        set_err("null array");
        return HFM_ERR_ARG;
    }
    const long N = s->g.N;
    int rc = check_finite(tfix, N, "tfix");
    if (rc) return rc;
    std::vector<uint8_t> fl(N);
    HIP_OK(hipStreamSynchronize(s->stream));
    HIP_OK(hipMemcpy(fl.data(), s->flags, N, hipMemcpyDeviceToHost));
    for (long t = 0; t < N; ++t) fl[t] = (uint8_t)((fl[t] & ~FLAG_FIXED) | (fixed[t] == 1 ? FLAG_FIXED : 0));
    HIP_OK(hipMemcpy(s->flags, fl.data(), N, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(s->tfix, tfix, N * sizeof(float), hipMemcpyHostToDevice));
    launch_apply_fixed(s, s->T);  // heat.py:192
    CHECK_LAUNCH();
    HIP_OK(hipStreamSynchronize(s->stream));
    return HFM_OK;
}

int hfm_set_source(hfm_scene* h, const int32_t* mask, double rate, double t_cap) {
    SCENE_OR_FAIL(h);
    if (!mask) {
        set_err("null array");
        return HFM_ERR_ARG;
    }
    if (!(rate == rate) || !(t_cap == t_cap)) {
        set_err("rate/t_cap not finite");
        return HFM_ERR_NONFINITE;
    }
    const long N = s->g.N;
    std::vector<uint8_t> fl(N);
    HIP_OK(hipStreamSynchronize(s->stream));
    HIP_OK(hipMemcpy(fl.data(), s->flags, N, hipMemcpyDeviceToHost));
    for (long t = 0; t < N; ++t) fl[t] = (uint8_t)((fl[t] & ~FLAG_SOURCE) | (mask[t] == 1 ? FLAG_SOURCE : 0));
    HIP_OK(hipMemcpy(s->flags, fl.data(), N, hipMemcpyHostToDevice));
    s->rate = rate;
    s->t_cap = t_cap;
    return HFM_OK;
}

int hfm_set_particles(hfm_scene* h, int64_t P, const float* pos, const float* vel, const float* temp) {
    SCENE_OR_FAIL(h);
    if (P < 0 || P > (int64_t)1 << 30 || (P > 0 && (!pos || !vel || !temp))) {
        set_err("bad particle arguments (P=%lld)", (long long)P);
        return HFM_ERR_ARG;
    }
    int rc = P > 0 ? check_finite(pos, 3 * P, "pos") : 0;
    if (!rc && P > 0) rc = check_finite(vel, 3 * P, "vel");
    if (!rc && P > 0) rc = check_finite(temp, P, "temp");
    if (rc) return rc;
    HIP_OK(hipStreamSynchronize(s->stream));
    if (P > s->P_cap) {
        void* olds[] = {s->pos, s->vel, s->temp, s->melt, s->yield, s->visc, s->pkey, s->pslot, s->sorted};
        for (void* q : olds)
            if (q) dfree(o, q);
        if ((rc = dalloc(o, &s->pos, 3 * P)) || (rc = dalloc(o, &s->vel, 3 * P)) || (rc = dalloc(o, &s->temp, P)) ||
            (rc = dalloc(o, &s->melt, P)) || (rc = dalloc(o, &s->yield, P)) || (rc = dalloc(o, &s->visc, P)) ||
            (rc = dalloc(o, &s->pkey, P)) || (rc = dalloc(o, &s->pslot, P)) || (rc = dalloc(o, &s->sorted, P)))
            return rc;
        s->P_cap = P;
    }
    s->P = P;
    if (P > 0) {
        HIP_OK(hipMemcpy(s->pos, pos, 3 * P * sizeof(float), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->vel, vel, 3 * P * sizeof(float), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->temp, temp, P * sizeof(float), hipMemcpyHostToDevice));
        std::vector<float> init(P, (float)s->prm.yield_solid);
        HIP_OK(hipMemcpy(s->yield, init.data(), P * sizeof(float), hipMemcpyHostToDevice));  // mpm.py:42
        std::fill(init.begin(), init.end(), (float)s->prm.mu_solid);
        HIP_OK(hipMemcpy(s->visc, init.data(), P * sizeof(float), hipMemcpyHostToDevice));   // mpm.py:43
        HIP_OK(hipMemset(s->melt, 0, P * sizeof(float)));                                     // solver.py:126
        s->bins_dirty = 1;
    } else {
        rc = reset_particle_fields(s);
        if (rc) return rc;
        HIP_OK(hipStreamSynchronize(s->stream));
    }
    return HFM_OK;
}

int hfm_update_particles(hfm_scene* h, const float* pos, const float* vel) {
    SCENE_OR_FAIL(h);
    if (s->P == 0) return HFM_OK;
    if (!pos || !vel) {
        set_err("null array");
        return HFM_ERR_ARG;
    }
    int rc = check_finite(pos, 3 * s->P, "pos");
    if (!rc) rc = check_finite(vel, 3 * s->P, "vel");
    if (rc) return rc;
    HIP_OK(hipMemcpyAsync(s->pos, pos, 3 * s->P * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_OK(hipMemcpyAsync(s->vel, vel, 3 * s->P * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    s->bins_dirty = 1;
    return HFM_OK;
}

int hfm_substep(hfm_scene* h, double dt) {
    SCENE_OR_FAIL(h);
    if (!(dt > 0) || !(dt < 1e3)) {
        set_err("bad dt %g", dt);
        return HFM_ERR_ARG;
    }
    return substep(o, dt);
}

int hfm_run(hfm_scene* h, double dt, int32_t K, double* ms) {
    SCENE_OR_FAIL(h);
    if (K < 0 || !(dt > 0)) {
        set_err("bad run arguments");
        return HFM_ERR_ARG;
    }
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    HIP_OK(hipEventRecord(e0, s->stream));
    for (int k = 0; k < K; ++k) {
        int rc = substep(o, dt);
        if (rc) {
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return rc;
        }
    }
    HIP_OK(hipEventRecord(e1, s->stream));
    HIP_OK(hipEventSynchronize(e1));
    float el = 0.f;
    HIP_OK(hipEventElapsedTime(&el, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (ms) *ms = el;
    return HFM_OK;
}

int64_t hfm_field_size(const hfm_scene* h, int32_t field) {
    if (!h) return -1;
    Scene* s = &owner(const_cast<hfm_scene*>(h))->s;
    if (field >= HFM_FIELD_AX && field <= HFM_FIELD_AZ) return s->g.F;
    long size = 0;
    if (!field_ptr(s, field, &size)) return -1;
    return size;
}

int hfm_get_field(hfm_scene* h, int32_t field, float* out) {
    SCENE_OR_FAIL(h);
    if (!out || field < 0 || field >= HFM_FIELD_COUNT) {
        set_err("bad field request %d", field);
        return HFM_ERR_ARG;
    }
    if (field >= HFM_FIELD_AX && field <= HFM_FIELD_AZ) {
        launch_face_fraction(s, field - HFM_FIELD_AX, s->scratch_field);
        CHECK_LAUNCH();
        HIP_OK(hipStreamSynchronize(s->stream));
        HIP_OK(hipMemcpy(out, s->scratch_field, s->g.F * sizeof(float), hipMemcpyDeviceToHost));
        return HFM_OK;
    }
    long size = 0;
    float* p = field_ptr(s, field, &size);
    HIP_OK(hipStreamSynchronize(s->stream));
    HIP_OK(hipMemcpy(out, p, size * sizeof(float), hipMemcpyDeviceToHost));
    return HFM_OK;
}

int hfm_set_field(hfm_scene* h, int32_t field, const float* in) {
    SCENE_OR_FAIL(h);
    if (!in || field < 0 || field >= HFM_FIELD_COUNT || (field >= HFM_FIELD_AX && field <= HFM_FIELD_AZ)) {
        set_err("field %d cannot be set", field);
        return HFM_ERR_ARG;
    }
    long size = 0;
    float* p = field_ptr(s, field, &size);
    int rc = check_finite(in, size, "field");
    if (rc) return rc;
    HIP_OK(hipStreamSynchronize(s->stream));
    HIP_OK(hipMemcpy(p, in, size * sizeof(float), hipMemcpyHostToDevice));
    return HFM_OK;
}

int hfm_get_cell_velocity(hfm_scene* h, float* out) {
    SCENE_OR_FAIL(h);
    if (!out) {
        set_err("null array");
        return HFM_ERR_ARG;
    }
    float* tmp = nullptr;
    HIP_OK(hipMalloc(&tmp, 3 * s->g.N * sizeof(float)));
    launch_cell_velocity(s, tmp);
    hipError_t e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) e = hipMemcpy(out, tmp, 3 * s->g.N * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(tmp);
    HIP_OK(e);
    return HFM_OK;
}

int hfm_get_particles(hfm_scene* h, float* temp, float* melt, float* yield_stress, float* viscosity) {
    SCENE_OR_FAIL(h);
    HIP_OK(hipStreamSynchronize(s->stream));
    if (s->P == 0) return HFM_OK;
    size_t b = s->P * sizeof(float);
    if (temp) HIP_OK(hipMemcpy(temp, s->temp, b, hipMemcpyDeviceToHost));
    if (melt) HIP_OK(hipMemcpy(melt, s->melt, b, hipMemcpyDeviceToHost));
    if (yield_stress) HIP_OK(hipMemcpy(yield_stress, s->yield, b, hipMemcpyDeviceToHost));
    if (viscosity) HIP_OK(hipMemcpy(viscosity, s->visc, b, hipMemcpyDeviceToHost));
    return HFM_OK;
}

int hfm_diagnostics(hfm_scene* h, double* divergence_max, double* speed_max) {
    SCENE_OR_FAIL(h);
    HIP_OK(hipMemsetAsync(s->umax, 0, 4 * sizeof(unsigned), s->stream));
    launch_divergence(s, (const float* const*)s->u, 1.0f, s->scratch_field, 0);
    launch_abs_max(s, s->scratch_field, s->g.N, &s->umax[0]);
    for (int a = 0; a < 3; ++a) launch_abs_max(s, s->u[a], s->g.F, &s->umax[1]);
    CHECK_LAUNCH();
    unsigned m[4];
    HIP_OK(hipStreamSynchronize(s->stream));
    HIP_OK(hipMemcpy(m, s->umax, sizeof(m), hipMemcpyDeviceToHost));
    float dm, sm;
    memcpy(&dm, &m[0], 4);
    memcpy(&sm, &m[1], 4);
    if (divergence_max) *divergence_max = (double)dm * s->prm.h;  // grid.py:260
    if (speed_max) *speed_max = (double)sm;                       // grid.py:263
    return HFM_OK;
}

int hfm_get_stats(hfm_scene* h, hfm_stats* out) {
    SCENE_OR_FAIL(h);
    if (!out) {
        set_err("null stats");
        return HFM_ERR_ARG;
    }
    *out = s->stats;
    return HFM_OK;
}

int hfm_kernel_time(hfm_scene* h, int32_t cls, double* total_ms, int64_t* launches) {
    SCENE_OR_FAIL(h);
    if (cls < 0 || cls >= HFM_KT_COUNT) {
        set_err("bad kernel class %d", cls);
        return HFM_ERR_ARG;
    }
    int rc = flush_timing(s);
    if (rc) return rc;
    if (total_ms) *total_ms = s->kt_ms[cls];
    if (launches) *launches = s->kt_n[cls];
    return HFM_OK;
}

int hfm_kernel_time_reset(hfm_scene* h) {
    SCENE_OR_FAIL(h);
    int rc = flush_timing(s);
    if (rc) return rc;
    for (int c = 0; c < HFM_KT_COUNT; ++c) {
        s->kt_ms[c] = 0.0;
        s->kt_n[c] = 0;
    }
    return HFM_OK;
}

int hfm_set_kernel_timing(hfm_scene* h, int32_t on) {
    SCENE_OR_FAIL(h);
    if (on) return enable_timing(s);
    int rc = flush_timing(s);
    s->timing = 0;
    return rc;
}

int hfm_synchronize(hfm_scene* h) {
    SCENE_OR_FAIL(h);
    HIP_OK(hipStreamSynchronize(s->stream));
    return HFM_OK;
}

}  // extern "C"
