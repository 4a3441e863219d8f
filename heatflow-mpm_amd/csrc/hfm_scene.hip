// hfm_scene.hip — C ABI of libhfm.so and the host-side order of the substep.
//
// HeatflowMPM.substep (solver.py:63-86) as launched here, for a batch of scenes on one stream
// (a single scene is a batch of one), one launch per stage over all scenes:
//   [particles] bin (when positions changed) + deterministic P2G gather
//   heat_prep (T2 = T) -> advect_T (into T) -> heat PCG (x = T) -> [particles] G2P
//   air_faces (u -> us) -> divergence -> numpy-exact rhs mean -> pressure PCG (warm start)
//   -> project (us -> u)
// The host never waits for the GPU inside a substep except to decide how many PCG iterations
// to enqueue: the first pass of each group of iterations posts the batch's progress to pinned
// host memory, every scene posts the end of its solve there, and the host keeps a bounded
// number of iterations in flight.
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <functional>
#include <cmath>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "hfm_device.h"
#include "hfm_scene.h"

using namespace hfm;

struct hfm_batch {
    Batch b;
};

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

#define RC(expr)                   \
    do {                           \
        int rc_ = (expr);          \
        if (rc_) return rc_;       \
    } while (0)

struct Owner {
    Scene s;
    std::vector<void*> dev;
    std::vector<size_t> dev_bytes;
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
    o->dev_bytes.push_back(bytes);
    HIP_OK(hipMemset(q, 0, bytes));
    *p = (T*)q;
    return HFM_OK;
}

void dfree(Owner* o, void* p) {
    for (size_t i = 0; i < o->dev.size(); ++i)
        if (o->dev[i] == p) {
            (void)hipFree(o->dev[i]);
            o->dev[i] = nullptr;
            o->dev_bytes[i] = 0;
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

// Serialises the persistent pressure solves of one device.
std::mutex& persist_mutex(int device) {
    static std::mutex mu[64];
    return mu[device & 63];
}

// Whether div_const's corrected product x * RN(1/b) + residual equals the IEEE quotient x / b for
// every significand (x in [1, 2)); with all intermediate values normal that covers every x with
// 2^-100 <= |x| <= 2^100 (the quotients scale exactly with x).  Results are cached per b.
bool const_division_exact(float b) {
    static std::mutex mu;
    static std::map<uint32_t, bool> cache;
    uint32_t key;
    memcpy(&key, &b, 4);
    {
        std::lock_guard<std::mutex> lock(mu);
        auto it = cache.find(key);
        if (it != cache.end()) return it->second;
    }
    bool exact = std::isfinite(b) && b != 0.0f && std::isfinite(1.0f / b);
    const float rb = 1.0f / b;
    for (uint32_t m = 0; exact && m < (1u << 23); ++m) {
        const uint32_t bits = 0x3f800000u | m;
        float x;
        memcpy(&x, &bits, 4);
        const float q = x * rb;
        const float e = std::fma(-q, b, x);
        const float r = std::fma(e, rb, q);
        exact = r == x / b;
    }
    std::lock_guard<std::mutex> lock(mu);
    cache[key] = exact;
    return exact;
}

// Scalar part of the descriptor, from the scene's parameters and state (fp32 as the reference
// passes them to its kernels).
void refresh_dev(Scene* s) {
    SceneDev& d = s->dev;
    const hfm_params& p = s->prm;
    d.n = s->n;
    d.N = s->N;
    d.F = s->F;
    d.h = (float)p.h;
    d.hh = d.h * d.h;
    d.rhh = 1.0f / d.hh;
    d.fastdiv = const_division_exact(d.hh) ? 1 : 0;
    d.rh = 1.0f / d.h;
    d.fast_h = const_division_exact(d.h) ? 1 : 0;
    d.rho = (float)p.rho;
    d.rho_d = p.rho;
    d.t_amb = (float)p.t_amb;
    d.g = (float)p.gravity;
    d.amin = (float)p.alpha_f_min;
    d.ka = (float)p.kappa_air;
    d.km = (float)p.kappa_mat;
    d.frac = (float)p.vol / (d.h * d.h * d.h);
    d.k0 = (float)p.k_particle;
    d.t_sol = (float)p.t_sol;
    d.t_liq = (float)p.t_liq;
    d.ys = (float)p.yield_solid;
    d.mus = (float)p.mu_solid;
    d.mul = (float)p.mu_liquid;
    d.t_max = (float)p.t_max;
    d.rate = (float)s->rate;
    d.t_cap = (float)s->t_cap;
    d.rate_on = s->rate > 0;
    d.buoyancy = p.buoyancy;
    s->version += 1;
}

// ---- batches -------------------------------------------------------------------------------

int batch_init(Batch* b, int device, hipStream_t stream, bool own_stream) {
    b->device = device;
    b->stream = stream;
    b->own_stream = own_stream;
    HIP_OK(hipHostMalloc((void**)&b->pace, sizeof(unsigned long long), hipHostMallocMapped | hipHostMallocCoherent));
    *b->pace = 0;
    HIP_OK(hipHostGetDevicePointer((void**)&b->pace_dev, b->pace, 0));
    return HFM_OK;
}

void batch_free_maps(Batch* b) {
    if (b->table) (void)hipFree(b->table);
    b->table = nullptr;
    b->table_cap = 0;
    for (int f = 0; f < FAM_COUNT; ++f) {
        if (b->maps[f]) (void)hipFree(b->maps[f]);
        b->maps[f] = nullptr;
        b->map_len[f] = 0;
        b->map_cap[f] = 0;
    }
    b->versions.clear();
}

// Device buffer of at least `count` elements, grown (never shrunk) as a batch changes.
template <class T>
int ensure_capacity(T** buf, size_t* cap, size_t count) {
    if (count <= *cap && *buf) return HFM_OK;
    if (*buf) HIP_OK(hipFree(*buf));
    *buf = nullptr;
    *cap = 0;
    const size_t want = std::max(count, (size_t)64) + count / 4;
    void* q = nullptr;
    HIP_OK(hipMalloc(&q, want * sizeof(T)));
    *buf = (T*)q;
    *cap = want;
    return HFM_OK;
}

void batch_release(Batch* b) {
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_free_maps(b);
    for (int r = 0; r < 3; ++r) {
        for (int f = 0; f < FAM_COUNT; ++f)
            if (b->fmap[r][f]) (void)hipFree(b->fmap[r][f]);
        if (b->fstage[r]) (void)hipHostFree(b->fstage[r]);
        b->fstage[r] = nullptr;
    }
    if (b->pmap) (void)hipFree(b->pmap);
    b->pmap = nullptr;
    b->pmap_cap = 0;
    if (b->pstage) (void)hipHostFree(b->pstage);
    b->pstage = nullptr;
    b->pstage_cap = 0;
    if (b->pace) (void)hipHostFree(b->pace);
    b->pace = nullptr;
    for (int i = 0; i < 2; ++i) {
        if (b->mg_exec[i]) (void)hipGraphExecDestroy(b->mg_exec[i]);
        b->mg_exec[i] = nullptr;
    }
    if (b->ev_pool) {
        for (int i = 0; i < b->ev_cap; ++i) (void)hipEventDestroy(b->ev_pool[i]);
        delete[] b->ev_pool;
        delete[] b->ev_class;
        b->ev_pool = nullptr;
    }
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    b->stream = nullptr;
}

// Workgroups scene s needs in sub-range sub of a launch family.
long family_blocks(const Scene* s, int f, int sub) {
    const SceneDev& d = s->dev;
    switch (f) {
        case FAM_CELL: return sub == 0 ? nblocks(s->N) : 0;
        case FAM_FACE3: return sub < 3 ? nblocks(s->F) : 0;
        case FAM_P2G: return d.P > 0 ? (sub == 0 ? nblocks(s->N) : nblocks(s->F)) : 0;
        case FAM_MEAN: return sub == 0 ? d.nchunks : 0;
        case FAM_PCG: return sub == 0 ? d.nb_upd : 0;
        case FAM_G2P: return sub == 0 && d.P > 0 ? nblocks(d.P) : 0;
        case FAM_DIR4: return sub == 0 && d.dtile == 4 && !d.cg_fused ? d.nb_dir : 0;
        case FAM_DIR8: return sub == 0 && d.dtile == 8 && !d.cg_fused ? d.nb_dir : 0;
        case FAM_DIR16: return sub == 0 && d.dtile == 16 && !d.cg_fused ? d.nb_dir : 0;
        case FAM_DIRFLAT: return sub == 0 && d.dtile == 0 ? d.nb_dir : 0;
        case FAM_FUSED4: return sub == 0 && d.dtile == 4 && d.cg_fused ? d.nb_dir : 0;
        case FAM_FUSED8: return sub == 0 && d.dtile == 8 && d.cg_fused ? d.nb_dir : 0;
        case FAM_FUSED16: return sub == 0 && d.dtile == 16 && d.cg_fused ? d.nb_dir : 0;
        case FAM_MGC: return sub == 0 && d.cg_mg ? 1 : 0;
        default:
            if (f >= FAM_MG0 && f <= FAM_MG7) {
                const int l = f - FAM_MG0;
                if (sub != 0 || !d.cg_mg || l >= d.mg_levels) return 0;
                const long m = d.mg_n[l];
                return nblocks(m * m * m);
            }
            return 0;
    }
}

}  // namespace

// Launch description of one family: the batched map, and for a batch of one its descriptor and
// grid (workgroups per sub-range x sub-ranges).
Launch hfm::family(const Batch* b, Family f) {
    Launch L{b->table, b->maps[f], b->map_len[f], b->stream, nullptr, dim3(0, 0, 1)};
    if (b->scenes.size() == 1) {
        const Scene* s = b->scenes[0];
        long x = 0;
        int y = 0;
        for (int sub = 0; sub < 4; ++sub) {
            const long nb = family_blocks(s, f, sub);
            if (nb > 0) {
                x = std::max(x, nb);
                y = sub + 1;
            }
        }
        L.solo = &s->dev;
        L.sgrid = dim3((unsigned)x, (unsigned)(y > 0 ? y : 1), 1);
        L.blocks = x > 0 ? (int)(x * y) : 0;
    }
    return L;
}

namespace {

// (Re)build the device descriptor table and block maps when a scene's descriptor changed.  Each
// (scene, sub-range) run of workgroups starts at a multiple of 8, so a workgroup's index inside
// its scene has the XCD of its hardware index and the XCD-aware tile mapping keeps working.
int batch_prepare(Batch* b) {
    const size_t ns = b->scenes.size();
    bool fresh = b->table && b->versions.size() == ns;
    for (size_t i = 0; fresh && i < ns; ++i) fresh = b->versions[i] == b->scenes[i]->version;
    if (fresh) return HFM_OK;
    HIP_OK(hipStreamSynchronize(b->stream));
    std::vector<SceneDev> tab(ns);
    for (size_t i = 0; i < ns; ++i) tab[i] = b->scenes[i]->dev;
    RC(ensure_capacity(&b->table, &b->table_cap, ns));
    HIP_OK(hipMemcpy(b->table, tab.data(), ns * sizeof(SceneDev), hipMemcpyHostToDevice));
    for (int f = 0; f < FAM_COUNT; ++f) {
        std::vector<MapEntry>& m = b->hmap[f];
        m.clear();
        b->ranges[f].assign(ns, std::make_pair((size_t)0, (size_t)0));
        for (size_t i = 0; i < ns; ++i) {
            b->ranges[f][i].first = m.size();
            for (int sub = 0; sub < 4; ++sub) {
                const long nb = family_blocks(b->scenes[i], f, sub);
                for (long l = 0; l < nb; ++l) m.push_back(map_entry((int)i, sub, l));
                while (m.size() % kXcds) m.push_back(kNoBlock);
            }
            b->ranges[f][i].second = m.size();
        }
        if (m.size() > (size_t)0x7fffffff) {
            set_err("batch too large: %zu workgroups in one launch", m.size());
            return HFM_ERR_ARG;
        }
        b->map_len[f] = (int)m.size();
        if (m.empty()) continue;
        RC(ensure_capacity(&b->maps[f], &b->map_cap[f], m.size()));
        HIP_OK(hipMemcpy(b->maps[f], m.data(), m.size() * sizeof(MapEntry), hipMemcpyHostToDevice));
    }
    b->versions.resize(ns);
    for (size_t i = 0; i < ns; ++i) b->versions[i] = b->scenes[i]->version;
    b->maps_epoch += 1;   // graphs captured over the old table and maps are stale
    return HFM_OK;
}

// Make b the batch that runs the work of its scenes: work a scene has queued on another stream
// is waited for first, so work on a scene stays ordered across batches.
int batch_adopt(Batch* b) {
    for (Scene* s : b->scenes)
        if (s->current != b) {
            if (s->current) HIP_OK(hipStreamSynchronize(s->current->stream));
            s->current = b;
        }
    return HFM_OK;
}

// A scene-level call: its own stream, after any batch work on the scene.
int scene_adopt(Scene* s) { return batch_adopt(s->solo); }

// ---- kernel timing -------------------------------------------------------------------------

int flush_timing(Batch* b) {
    if (!b->timing || b->ev_used == 0) return HFM_OK;
    HIP_OK(hipStreamSynchronize(b->stream));
    for (int e = 0; e + 1 < b->ev_used; e += 2) {
        float ms = 0.f;
        HIP_OK(hipEventElapsedTime(&ms, b->ev_pool[e], b->ev_pool[e + 1]));
        b->kt_ms[b->ev_class[e]] += ms;
        b->kt_n[b->ev_class[e]] += 1;
    }
    b->ev_used = 0;
    return HFM_OK;
}

// A pair of events for one timed launch of kernel class cls (flushing the pool when full).
int timing_events(Batch* b, int cls, hipEvent_t* ev0, hipEvent_t* ev1) {
    *ev0 = *ev1 = nullptr;
    if (!b->timing) return HFM_OK;
    if (b->ev_used + 2 > b->ev_cap) RC(flush_timing(b));
    b->ev_class[b->ev_used] = cls;
    *ev0 = b->ev_pool[b->ev_used++];
    *ev1 = b->ev_pool[b->ev_used++];
    return HFM_OK;
}

int enable_timing(Batch* b) {
    if (!b->ev_pool) {
        b->ev_cap = 4096;
        b->ev_pool = new hipEvent_t[b->ev_cap];
        b->ev_class = new int[b->ev_cap];
        for (int i = 0; i < b->ev_cap; ++i) HIP_OK(hipEventCreate(&b->ev_pool[i]));
    }
    b->timing = 1;
    return HFM_OK;
}

int set_timing(Batch* b, int on) {
    if (on) return enable_timing(b);
    int rc = flush_timing(b);
    b->timing = 0;
    return rc;
}

int kernel_time(Batch* b, int cls, double* total_ms, int64_t* launches) {
    if (cls < 0 || cls >= HFM_KT_COUNT) {
        set_err("bad kernel class %d", cls);
        return HFM_ERR_ARG;
    }
    RC(flush_timing(b));
    if (total_ms) *total_ms = b->kt_ms[cls];
    if (launches) *launches = b->kt_n[cls];
    return HFM_OK;
}

int kernel_time_reset(Batch* b) {
    RC(flush_timing(b));
    for (int c = 0; c < HFM_KT_COUNT; ++c) {
        b->kt_ms[c] = 0.0;
        b->kt_n[c] = 0;
    }
    return HFM_OK;
}

// ---- PCG driver ----------------------------------------------------------------------------

// Launches of families fams[0..nf) restricted to the scenes with active[i] set: the scenes' runs
// of the host maps are concatenated in a pinned staging buffer and copied on the batch stream
// into the next of three rotating device buffers.
int filter_families(Batch* b, const Family* fams, int nf, const std::vector<char>& active, Launch* out) {
    const int r = b->fslot;
    b->fslot = (b->fslot + 1) % 3;
    size_t total = 0;
    for (int q = 0; q < nf; ++q) total += b->hmap[fams[q]].size();
    if (total > b->fstage_cap[r]) {
        if (b->fstage[r]) HIP_OK(hipHostFree(b->fstage[r]));
        b->fstage[r] = nullptr;
        b->fstage_cap[r] = 0;
        HIP_OK(hipHostMalloc((void**)&b->fstage[r], total * sizeof(MapEntry), hipHostMallocDefault));
        b->fstage_cap[r] = total;
    }
    size_t off = 0;
    for (int q = 0; q < nf; ++q) {
        const Family f = fams[q];
        MapEntry* dst = b->fstage[r] + off;
        size_t len = 0;
        for (size_t i = 0; i < b->scenes.size(); ++i) {
            if (!active[i]) continue;
            const auto rg = b->ranges[f][i];
            memcpy(dst + len, b->hmap[f].data() + rg.first, (rg.second - rg.first) * sizeof(MapEntry));
            len += rg.second - rg.first;
        }
        Launch L = family(b, f);
        L.blocks = (int)len;
        if (len > 0) {
            RC(ensure_capacity(&b->fmap[r][f], &b->fmap_cap[r][f], len));
            HIP_OK(hipMemcpyAsync(b->fmap[r][f], dst, len * sizeof(MapEntry), hipMemcpyHostToDevice, b->stream));
            L.map = b->fmap[r][f];
        }
        out[f] = L;
        off += len;
    }
    return HFM_OK;
}

// Most multigrid levels (tail: the deepest tail level) among the batch's scenes with the
// multigrid-preconditioned pressure PCG (0: none).
int mg_depth(const Batch* b, bool tail) {
    int L = 0;
    for (const Scene* s : b->scenes)
        if (s->dev.cg_mg) L = std::max(L, tail ? s->dev.mg_tail : s->dev.mg_levels);
    return L;
}

// The multigrid coarse levels depend on the solid fraction only: formed again when any multigrid
// scene of the batch has a new one (for every scene of the batch; unchanged ones get the same values).
void launch_mg_coefficients_if_stale(Batch* b) {
    bool stale = false;
    for (Scene* s : b->scenes) stale |= s->dev.cg_mg && s->mg_coef_dirty;
    if (!stale) return;
    launch_mg_coefficients(b, mg_depth(b, false));
    for (Scene* s : b->scenes) s->mg_coef_dirty = 0;
}

// The pressure solves of a batch's scenes marked in `persist` (each one's tiles fit the device at once) as
// persistent launches: scenes of one tile depth packed first-fit, largest first, into cooperative
// launches of at most the device's resident workgroups, each scene's block-map run whole (its padding to
// the XCD stride included).  Every scene's workgroups wait only on each other, so the scenes of a group
// solve independently and a group's launch ends with its slowest solve.  All groups' maps go to the device
// in one copy ahead of the launches; the staging buffer is rewritten only by the next solve, after this
// one has ended.
int launch_persist_groups(Batch* b, const std::vector<char>& persist, bool timed) {
    const size_t ns = b->scenes.size();
    struct Group {
        int D;
        std::vector<size_t> scenes;
        size_t len = 0;
    };
    std::vector<Group> groups;
    for (int D : {16, 8, 4}) {
        const Family f = D == 16 ? FAM_FUSED16 : (D == 8 ? FAM_FUSED8 : FAM_FUSED4);
        const int cap = pcg_persist_capacity(D, b->device, true);
        std::vector<size_t> idx;
        for (size_t i = 0; i < ns; ++i)
            if (persist[i] && b->scenes[i]->dev.dtile == D) idx.push_back(i);
        std::stable_sort(idx.begin(), idx.end(), [&](size_t x, size_t y) {
            const auto rx = b->ranges[f][x], ry = b->ranges[f][y];
            return rx.second - rx.first > ry.second - ry.first;
        });
        const size_t first = groups.size();
        for (size_t i : idx) {
            const size_t len = b->ranges[f][i].second - b->ranges[f][i].first;
            if (len > (size_t)cap) {
                set_err("persistent pressure PCG: scene %zu needs %zu workgroups, the device holds %d", i, len, cap);
                return HFM_ERR_STATE;
            }
            size_t g = first;
            while (g < groups.size() && groups[g].len + len > (size_t)cap) ++g;
            if (g == groups.size()) {
                groups.push_back(Group());
                groups.back().D = D;
            }
            groups[g].scenes.push_back(i);
            groups[g].len += len;
        }
    }
    if (groups.empty()) return HFM_OK;
    size_t total = 0;
    for (const Group& g : groups) total += g.len;
    if (total > b->pstage_cap) {
        if (b->pstage) HIP_OK(hipHostFree(b->pstage));
        b->pstage = nullptr;
        b->pstage_cap = 0;
        HIP_OK(hipHostMalloc((void**)&b->pstage, total * sizeof(MapEntry), hipHostMallocDefault));
        b->pstage_cap = total;
    }
    size_t off = 0;
    for (const Group& g : groups) {
        const Family f = g.D == 16 ? FAM_FUSED16 : (g.D == 8 ? FAM_FUSED8 : FAM_FUSED4);
        for (size_t i : g.scenes) {
            const auto rg = b->ranges[f][i];
            memcpy(b->pstage + off, b->hmap[f].data() + rg.first, (rg.second - rg.first) * sizeof(MapEntry));
            off += rg.second - rg.first;
        }
    }
    RC(ensure_capacity(&b->pmap, &b->pmap_cap, total));
    HIP_OK(hipMemcpyAsync(b->pmap, b->pstage, total * sizeof(MapEntry), hipMemcpyHostToDevice, b->stream));
    off = 0;
    for (const Group& g : groups) {
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timed) RC(timing_events(b, HFM_KT_PCG_PERSIST, &e0, &e1));
        const hipError_t e = launch_pcg_persist_batch(b->table, b->pmap + off, (int)g.len, g.D, b->stream, e0, e1);
        if (e != hipSuccess) {
            set_err("persistent pressure PCG: cooperative launch of %zu workgroups (%zu scenes) failed: %s", g.len,
                    g.scenes.size(), hipGetErrorString(e));
            return HFM_ERR_HIP;
        }
        off += g.len;
    }
    return HFM_OK;
}

// Run PCG iterations from iteration 0 for every scene of the batch whose mirror is not done
// (the init launches are enqueued), and return when the device reports all of them done.
// ir: the inner solves of a mixed-precision refinement step (FP16 kernels).
// tail: launches that follow the solve in the stream; where the whole solve is already enqueued (the
// persistent form) they are enqueued before the wait for it, so the GPU does not idle while the host
// learns that the solve has ended; returns whether it did (else the caller enqueues them)
int pcg_iterate(Batch* b, int op, float dt, bool ir, int maxeff, int pred, const std::function<void()>& tail,
                bool* tail_done) {
    const unsigned tag = ++b->solve_tag;
    const int mg = op == 1 ? mg_depth(b, true) : 0;
    // pass A launches: the heat operator uses the flat kernel with the PCG geometry; the
    // pressure operator one launch per tiling present in the batch, two-pass or fused
    Family dirs[8];
    int ndir = 0;
    bool two_pass = op == 0;
    if (op == 0) {
        dirs[ndir++] = FAM_PCG;
    } else {
        for (Family f : {FAM_DIR4, FAM_DIR8, FAM_DIR16, FAM_DIRFLAT, FAM_FUSED4, FAM_FUSED8, FAM_FUSED16})
            if (b->map_len[f] > 0) dirs[ndir++] = f;
        bool any_fused = false;
        for (Scene* s : b->scenes) {
            if (s->dev.cg_fused)
                any_fused = true;
            else
                two_pass = true;
        }
        (void)any_fused;
    }
    auto is_fused = [](Family f) { return f == FAM_FUSED4 || f == FAM_FUSED8 || f == FAM_FUSED16; };
    // Launches of the iteration families; in a batch they are restricted, between groups of
    // iterations, to the scenes whose solve still runs.
    Launch fl[FAM_COUNT];
    for (int f = 0; f < FAM_COUNT; ++f) fl[f] = family(b, (Family)f);
    Family iter_fams[8];
    int niter = 0;
    for (int d = 0; d < ndir; ++d) iter_fams[niter++] = dirs[d];
    if (op == 1 && two_pass) iter_fams[niter++] = FAM_PCG;  // pressure pass B (heat: already dirs[0])
    const size_t ns = b->scenes.size();
    std::vector<char> active(ns, 1);
    size_t n_active = ns;
    if (ir && ns > 1) {   // scenes whose solve has ended take no part in this inner solve
        bool changed = false;
        for (size_t i = 0; i < ns; ++i)
            if (__atomic_load_n(&b->scenes[i]->mirror[op].done, __ATOMIC_ACQUIRE)) {
                active[i] = 0;
                --n_active;
                changed = true;
            }
        if (changed && n_active > 0) RC(filter_families(b, iter_fams, niter, active, fl));
    }
    auto refresh_active = [&]() -> int {
        if (ns < 2) return HFM_OK;
        bool changed = false;
        for (size_t i = 0; i < ns; ++i)
            if (active[i] && __atomic_load_n(&b->scenes[i]->mirror[op].done, __ATOMIC_ACQUIRE)) {
                active[i] = 0;
                --n_active;
                changed = true;
            }
        if (changed && n_active > 0) return filter_families(b, iter_fams, niter, active, fl);
        return HFM_OK;
    };
    // One launch of pass A (or of the fused iteration) per family for iteration k; the first
    // non-empty launch posts the progress.
    auto pass_a = [&](int k, unsigned long long* pace, bool timed) -> int {
        unsigned long long* pc = pace;
        for (int d = 0; d < ndir; ++d) {
            const Launch& L = fl[dirs[d]];
            if (L.blocks <= 0) continue;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (is_fused(dirs[d])) {
                if (timed) RC(timing_events(b, HFM_KT_PCG_FUSED, &e0, &e1));
                launch_pcg_fused(L, dirs[d], k, pc, tag, e0, e1);
            } else {
                if (timed) RC(timing_events(b, HFM_KT_PCG_DIR, &e0, &e1));
                launch_pcg_dir(L, dirs[d], op, dt, k, pc, tag, e0, e1, ir);
            }
            pc = nullptr;
        }
        return HFM_OK;
    };
    // Iterations are enqueued in groups, at most two groups in flight.  Each iteration is a
    // pass A and a pass B (two-pass form) or one fused launch; pass A of iteration k+1 finishes
    // iteration k, so the first launch of a group posts how many iterations the previous groups
    // completed, which releases the next group; after the last iteration one more pass A closes
    // every solve.  Group size follows the iteration count of the previous solve so that few
    // launches are wasted after convergence.
    int launched = 0;
    bool past_pred = false;
    long cells = 0;
    for (Scene* s : b->scenes) cells += (long)s->n * s->n * s->n;
    const bool single_ok = ir || cells >= 96L * 96 * 96;
    int group_end[2] = {0, 0};
    int inflight = 0;
    bool closed = false;
    const auto t0 = std::chrono::steady_clock::now();
    long spins = 0;
    const bool timed = op == 1 && b->timing;
    // The persistent form runs the whole solve in one cooperative launch; the loop below then only
    // waits for it.  One such launch at a time per device: two of them, each partly resident, would
    // wait on each other until their spin limits.
    std::unique_lock<std::mutex> persist_lock;
    if (op == 1 && !ir && ns == 1 && ndir == 1 && b->scenes[0]->dev.cg_persist && fl[dirs[0]].solo &&
        fl[dirs[0]].blocks > 0) {
        persist_lock = std::unique_lock<std::mutex>(persist_mutex(b->device));
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (timed) RC(timing_events(b, HFM_KT_PCG_PERSIST, &e0, &e1));
        const hipError_t e = launch_pcg_persist(fl[dirs[0]], b->scenes[0]->dev.dtile, e0, e1);
        if (e != hipSuccess) {
            set_err("persistent pressure PCG: cooperative launch of %d workgroups failed: %s", fl[dirs[0]].blocks,
                    hipGetErrorString(e));
            return HFM_ERR_HIP;
        }
        launched = maxeff;
        closed = true;
        if (tail) {
            tail();
            CHECK_LAUNCH();
            *tail_done = true;
        }
    } else if (op == 1 && !ir && ns > 1) {
        // a batch: the scenes that can run the persistent form solve in cooperative groups first, the
        // others (if any) in the loop below, from which the persistent ones are left out
        std::vector<char> persist(ns, 0);
        size_t np = 0;
        for (size_t i = 0; i < ns; ++i)
            if (active[i] && b->scenes[i]->dev.cg_persist) {
                persist[i] = 1;
                ++np;
            }
        if (np > 0) {
            persist_lock = std::unique_lock<std::mutex>(persist_mutex(b->device));
            RC(launch_persist_groups(b, persist, timed));
            CHECK_LAUNCH();
            for (size_t i = 0; i < ns; ++i)
                if (persist[i]) {
                    active[i] = 0;
                    --n_active;
                }
            if (n_active == 0) {
                launched = maxeff;
                closed = true;
                if (tail) {
                    tail();
                    CHECK_LAUNCH();
                    *tail_done = true;
                }
            } else {
                RC(filter_families(b, iter_fams, niter, active, fl));
            }
        }
    }
    for (;;) {
        bool all_done = true;
        for (Scene* s : b->scenes)
            if (!__atomic_load_n(&s->mirror[op].done, __ATOMIC_ACQUIRE)) {
                all_done = false;
                break;
            }
        if (all_done) break;
        const unsigned long long w = __atomic_load_n(b->pace, __ATOMIC_ACQUIRE);
        const int prog = (unsigned)(w >> 32) == tag ? (int)(w & 0xffffffffu) : 0;
        while (inflight > 0 && prog >= group_end[0]) {
            group_end[0] = group_end[1];
            --inflight;
        }
        if (inflight < 2 && launched < maxeff) {
            RC(refresh_active());
            // groups past the predicted count hold one iteration (the next group's first pass A finishes
            // the last iteration, so a solve that ends as predicted leaves one queued, not two): the
            // first of them always; every one for short solves (a prediction of at most 4: the heat
            // solve, the refinement's inner solves) whose iterations outlast the host's turnaround (an
            // inner solve's, with its V-cycle, or a batch of at least 96^3 cells); otherwise groups of
            // two, so that the host does not pace short iterations one by one
            int g = pred - launched;
            if (launched >= pred) {
                g = (pred <= 4 && single_ok) || !past_pred ? 1 : 2;
                past_pred = true;
            } else {
                g = g < 2 ? 2 : (g > 16 ? 16 : g);
            }
            if (g > maxeff - launched) g = maxeff - launched;
            for (int q = 0; q < g; ++q) {
                const int k = launched + q;
                RC(pass_a(k, q == 0 ? b->pace_dev : nullptr, timed));
                if (two_pass) {
                    hipEvent_t e0 = nullptr, e1 = nullptr;
                    if (timed) RC(timing_events(b, HFM_KT_PCG_UPD, &e0, &e1));
                    launch_pcg_upd(fl[FAM_PCG], op, k, e0, e1, ir);
                    if (mg) RC(launch_mg_vcycle(b, mg, ir));
                }
            }
            CHECK_LAUNCH();
            launched += g;
            group_end[inflight++] = launched;
            continue;
        }
        if (launched >= maxeff && !closed) {
            RC(refresh_active());
            RC(pass_a(maxeff, b->pace_dev, false));
            CHECK_LAUNCH();
            closed = true;
            continue;
        }
        if ((++spins & 0x3ff) == 0) {
            const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (el > 60.0) {
                set_err("PCG (%s) did not finish within 60 s (progress %d, %d iterations launched)",
                        op ? "pressure" : "heat", prog, launched);
                return HFM_ERR_HIP;
            }
            hipError_t e = hipStreamQuery(b->stream);
            if (e != hipSuccess && e != hipErrorNotReady) {
                set_err("HIP error seen while waiting for the %s PCG (substep %lld of the %s, iteration %d "
                        "reached): %s.  The stream reports the first fault of any kernel enqueued before this "
                        "poll, so the faulting kernel may be any earlier kernel of this substep, not the PCG",
                        op ? "pressure" : "heat", (long long)b->scenes[0]->stats.substeps + 1,
                        ns > 1 ? "batch" : "scene", prog, hipGetErrorString(e));
                return HFM_ERR_HIP;
            }
            std::this_thread::yield();
        }
    }
    return HFM_OK;
}

// Enqueue one PCG solve (op 0 heat, 1 pressure) of every scene of the batch and return when
// the device reports all of them done.
int run_pcg(Batch* b, int op, float dt, const std::function<void()>& tail, bool* tail_done) {
    int maxeff = 0, pred = 2;
    bool any_fused = false;
    for (Scene* s : b->scenes) {
        CgMirror* m = &s->mirror[op];
        __atomic_store_n(&m->it, 0, __ATOMIC_RELAXED);
        __atomic_store_n(&m->done, 0, __ATOMIC_RELEASE);
        const int mi = op ? s->prm.cg_max_iter : s->prm.heat_max_iter;
        maxeff = std::max(maxeff, mi + (mi & 1));
        pred = std::max(pred, op ? s->pres_pred : s->heat_pred);
        if (op == 1 && s->dev.cg_fused) any_fused = true;
    }
    launch_pcg_init(family(b, FAM_PCG), op, dt);
    if (any_fused) launch_pcg_init_fused(family(b, FAM_PCG));
    if (op == 1 && mg_depth(b, false)) {   // multigrid: the coarse levels (when the solid fraction changed), z = M r
        launch_mg_coefficients_if_stale(b);
        RC(launch_mg_vcycle(b, mg_depth(b, true)));
    }
    CHECK_LAUNCH();
    RC(pcg_iterate(b, op, dt, false, maxeff, pred, tail, tail_done));
    int bad = -1;
    for (size_t si = 0; si < b->scenes.size(); ++si) {
        Scene* s = b->scenes[si];
        CgMirror* m = &s->mirror[op];
        const int it = __atomic_load_n(&m->it, __ATOMIC_ACQUIRE);
        const int conv = __atomic_load_n(&m->converged, __ATOMIC_ACQUIRE);
        if (conv == -2) {
            set_err("persistent pressure PCG: a workgroup's wait for the others ran out of polls at iteration %d "
                    "(every workgroup has ended; the pressure is not solved)", it);
            return HFM_ERR_HIP;
        }
        // a breakdown stop keeps the last iterate; a non-finite residual means the state already is
        if (conv < 0 && !(m->rr <= 1.0e300) && bad < 0) bad = (int)si;
        const int maxiter = op ? s->prm.cg_max_iter : s->prm.heat_max_iter;
        const int ce = op ? s->prm.cg_check_every : s->prm.heat_check_every;
        const int reported = it == 0 && conv ? 0 : warp_reported_count(it, conv, maxiter, ce);
        if (op == 1) {
            s->pres_pred = it > 2 ? it : 2;
            s->stats.p_iters_total += it;
            s->stats.p_iters_exact = it;
            s->stats.p_iters = reported;
            s->stats.p_converged = conv;
            s->stats.p_rr = m->rr;
            s->stats.p_atol = m->atol;
            s->stats.rhs_mean = m->mean;
        } else {
            s->heat_pred = it > 2 ? it : 2;
            s->stats.source_dT_sum = s->dev.rate_on ? m->src_dt : 0.0;
            s->stats.h_iters_total += it;
            s->stats.h_iters_exact = it;
            s->stats.h_iters = reported;
            s->stats.h_converged = conv;
            s->stats.h_rr = m->rr;
            s->stats.h_atol = m->atol;
        }
    }
    if (bad >= 0) {
        set_err("%s PCG of %s %d stopped on a non-finite residual: the scene's state is not finite",
                op ? "pressure" : "heat", b->scenes.size() > 1 ? "batch member" : "scene", bad);
        return HFM_ERR_NONFINITE;
    }
    return HFM_OK;
}

// One solve of every (mixed-precision) scene of the batch as iterative refinement: the true
// residual and its scaling (k_ir_residual, k_ir_scale), then an inner FP16 PCG, repeated until
// the scale kernel ends each scene's solve.  The host learns of both kinds of end through the
// mirror's `done`; `end` tells them apart.
int run_ir(Batch* b, int op, float dt) {
    int maxeff = 0, pred = 2;
    for (Scene* s : b->scenes) {
        CgMirror* m = &s->mirror[op];
        __atomic_store_n(&m->end, 0, __ATOMIC_RELAXED);
        __atomic_store_n(&m->done, 0, __ATOMIC_RELEASE);
        maxeff = std::max(maxeff, op ? s->prm.cg_max_iter : s->prm.heat_max_iter);
        pred = std::max(pred, op ? s->pres_pred : s->heat_pred);
    }
    const Launch fam = family(b, FAM_PCG);
    // multigrid scenes: this solve's coarse levels, and z = M r after every scale step (the inner
    // solves' init)
    const int mg = op == 1 ? mg_depth(b, true) : 0;
    if (mg) launch_mg_coefficients_if_stale(b);
    launch_ir_residual(fam, op, dt, true);
    launch_ir_scale(fam, op, 0);
    if (mg) RC(launch_mg_vcycle(b, mg, true));
    CHECK_LAUNCH();
    for (int step = 0;; ++step) {
        RC(pcg_iterate(b, op, dt, true, maxeff, pred, {}, nullptr));
        bool all_end = true;
        int longest = 2;
        for (Scene* s : b->scenes) {
            CgMirror* m = &s->mirror[op];
            if (__atomic_load_n(&m->end, __ATOMIC_ACQUIRE)) continue;
            all_end = false;
            longest = std::max(longest, __atomic_load_n(&m->it, __ATOMIC_ACQUIRE));
            __atomic_store_n(&m->done, 0, __ATOMIC_RELEASE);  // before the launches that may set it
        }
        if (all_end) break;
        pred = longest;
        if (step > 4 * maxeff + 8) {
            set_err("%s refinement did not end", op ? "pressure" : "heat");
            return HFM_ERR_STATE;
        }
        launch_ir_residual(fam, op, dt, false);
        launch_ir_scale(fam, op, step + 1);
        if (mg) RC(launch_mg_vcycle(b, mg, true));
        CHECK_LAUNCH();
    }
    for (size_t si = 0; si < b->scenes.size(); ++si) {
        Scene* s = b->scenes[si];
        CgMirror* m = &s->mirror[op];
        const int it = __atomic_load_n(&m->it, __ATOMIC_ACQUIRE);
        const int conv = __atomic_load_n(&m->converged, __ATOMIC_ACQUIRE);
        const int outer = __atomic_load_n(&m->outer, __ATOMIC_ACQUIRE);
        if (conv < 0) {
            set_err("%s solve of %s %zu (n = %d) in substep %lld: non-finite residual after %d inner iterations "
                    "in %d refinement steps", op ? "pressure" : "heat", b->scenes.size() > 1 ? "batch member" : "scene",
                    si, s->n, (long long)s->stats.substeps + 1, it, outer);
            return HFM_ERR_NONFINITE;
        }
        const int inner = outer > 0 ? (it + outer - 1) / outer : 2;
        if (op == 1) {
            s->pres_pred = inner > 2 ? inner : 2;
            s->stats.p_iters_total += it;
            s->stats.p_iters_exact = it;
            s->stats.p_iters = it;
            s->stats.p_converged = conv;
            s->stats.p_rr = m->rr;
            s->stats.p_atol = m->atol;
            s->stats.p_outer = outer;
            s->stats.rhs_mean = m->mean;
        } else {
            s->heat_pred = inner > 2 ? inner : 2;
            s->stats.source_dT_sum = s->dev.rate_on ? m->src_dt : 0.0;
            s->stats.h_iters_total += it;
            s->stats.h_iters_exact = it;
            s->stats.h_iters = it;
            s->stats.h_converged = conv;
            s->stats.h_rr = m->rr;
            s->stats.h_atol = m->atol;
            s->stats.h_outer = outer;
        }
    }
    return HFM_OK;
}

int batch_substep(Batch* b, double dt) {
    RC(batch_prepare(b));
    RC(batch_adopt(b));
    const float dt32 = (float)dt;
    // particle-grid gather: every lattice after a re-binning; while the particles stay put, the cell
    // lattice is counted once more by the general gather, its contribution list built from the
    // counts, and from then on summed from the list (host and device step this state together:
    // k_bin_pack marks the list stale, k_csr_fill ready)
    bool general = false, listed = false;
    for (Scene* s : b->scenes) {
        if (s->dev.P == 0) continue;
        if (s->bins_dirty) {
            launch_bin_particles(s, b->stream);
            s->csr_state = 0;
            s->mg_coef_dirty = 1;   // this substep's gather forms a new solid fraction
        } else if (s->dev.P <= kCsrMaxParticles && 8L * s->dev.P <= s->csr_cap) {   // 8 P int-indexed terms
            s->csr_state = s->csr_state == 0 ? 1 : 2;
        }
        general |= s->csr_state < 2;
        listed |= s->csr_state == 2;
    }
    if (general) launch_p2g(family(b, FAM_P2G));
    for (Scene* s : b->scenes)
        if (s->dev.P > 0 && s->csr_state == 1) launch_csr_build(s, b->stream);
    if (listed) launch_p2g_listed(family(b, FAM_CELL));
    const bool mixed = !b->scenes.empty() && b->scenes[0]->dev.mixed != 0;  // one precision per batch
    launch_heat_prep(family(b, FAM_CELL), dt32);
    launch_advect_T(family(b, FAM_CELL), dt32);
    bool mc = false;
    for (Scene* s : b->scenes) mc |= s->dev.maccormack != 0;
    if (mc) launch_mc_T(family(b, FAM_CELL), dt32);
    CHECK_LAUNCH();
    RC(mixed ? run_ir(b, 0, dt32) : run_pcg(b, 0, dt32, {}, nullptr));
    launch_g2p(family(b, FAM_G2P));
    launch_air_faces(family(b, FAM_FACE3), dt32);
    if (mc) launch_mc_faces(family(b, FAM_FACE3), dt32);
    launch_divergence(family(b, FAM_CELL), dt);
    launch_rhs_mean(family(b, FAM_MEAN));
    CHECK_LAUNCH();
    bool projected = false;   // the projection enqueued behind a persistent solve, before the wait for it
    const auto project = [&] { launch_project(family(b, FAM_FACE3), dt); };
    RC(mixed ? run_ir(b, 1, dt32) : run_pcg(b, 1, dt32, project, &projected));
    if (!projected) launch_project(family(b, FAM_FACE3), dt);
    CHECK_LAUNCH();
    for (Scene* s : b->scenes) s->stats.substeps += 1;
    return HFM_OK;
}

int batch_run(Batch* b, double dt, int K, double* ms) {
    RC(batch_prepare(b));
    RC(batch_adopt(b));
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    int rc = HFM_OK;
    if (hipEventRecord(e0, b->stream) != hipSuccess) rc = HFM_ERR_HIP;
    for (int k = 0; k < K && !rc; ++k) rc = batch_substep(b, dt);
    float el = 0.f;
    if (!rc) {
        hipError_t e = hipEventRecord(e1, b->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        if (e == hipSuccess) e = hipEventElapsedTime(&el, e0, e1);
        if (e != hipSuccess) {
            set_err("run timing: %s", hipGetErrorString(e));
            rc = HFM_ERR_HIP;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (!rc && ms) *ms = el;
    return rc;
}

float* field_ptr(Scene* s, int f, long* size) {
    const long N = s->N, F = s->F;
    SceneDev& d = s->dev;
    switch (f) {
        case HFM_FIELD_T: *size = N; return d.T;
        case HFM_FIELD_T2: *size = N; return d.T2;
        case HFM_FIELD_TW: *size = N; return d.Tw;
        case HFM_FIELD_U: *size = F; return d.u[0];
        case HFM_FIELD_V: *size = F; return d.u[1];
        case HFM_FIELD_W: *size = F; return d.u[2];
        case HFM_FIELD_P: *size = N; return d.p;
        case HFM_FIELD_RHS: *size = N; return d.rhs;
        case HFM_FIELD_ALPHA_S: *size = N; return d.alpha_s;
        case HFM_FIELD_ALPHA_P: *size = N; return d.alpha_p;
        case HFM_FIELD_TSUM: *size = N; return d.tsum;
        case HFM_FIELD_KAPPA: *size = N; return d.kap;
        case HFM_FIELD_QX: case HFM_FIELD_QY: case HFM_FIELD_QZ: *size = F; return d.q[f - HFM_FIELD_QX];
        case HFM_FIELD_KX: case HFM_FIELD_KY: case HFM_FIELD_KZ: *size = F; return d.kf[f - HFM_FIELD_KX];
        case HFM_FIELD_MX: case HFM_FIELD_MY: case HFM_FIELD_MZ: *size = F; return d.mf[f - HFM_FIELD_MX];
        case HFM_FIELD_PDIAG: *size = N; return d.pdiag;
        case HFM_FIELD_HRHS: *size = N; return d.hrhs;
        case HFM_FIELD_HDIAG: *size = N; return d.hdiag;
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
    const long N = s->N, F = s->F;
    SceneDev& d = s->dev;
    HIP_OK(hipMemsetAsync(d.alpha_p, 0, N * sizeof(float), s->stream));
    HIP_OK(hipMemsetAsync(d.tsum, 0, N * sizeof(float), s->stream));
    HIP_OK(hipMemcpyAsync(d.alpha_s, d.st_alpha, N * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
    s->mg_coef_dirty = 1;
    for (int a = 0; a < 3; ++a) {
        HIP_OK(hipMemsetAsync(d.q[a], 0, F * sizeof(float), s->stream));
        HIP_OK(hipMemcpyAsync(d.kf[a], d.st_k[a], F * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
        HIP_OK(hipMemcpyAsync(d.mf[a], d.st_m[a], F * sizeof(float), hipMemcpyDeviceToDevice, s->stream));
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
    (void)o;                                      \
    HIP_OK(hipSetDevice(s->device));              \
    RC(scene_adopt(s));

#define BATCH_OR_FAIL(h)                          \
    if (!(h)) {                                   \
        set_err("null batch");                    \
        return HFM_ERR_ARG;                       \
    }                                             \
    Batch* b = &(h)->b;                           \
    HIP_OK(hipSetDevice(b->device));

}  // namespace

void hfm::set_error(const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

int hfm::check_finite_host(const float* a, long n, const char* what) { return check_finite(a, n, what); }

extern "C" {

const char* hfm_last_error(void) { return g_err.c_str(); }

int hfm_device_count(void) {
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) return 0;
    return c;
}

int hfm_mesh_inside(const float* tri, int64_t nt, const float* points, int64_t np, float threshold, int32_t* flags,
                    float* winding, int device) {
    if (nt < 0 || np < 0 || nt > ((int64_t)1 << 26) || np > ((int64_t)1 << 30) || (nt > 0 && !tri) ||
        (np > 0 && (!points || !flags)) || !(threshold == threshold)) {
        set_err("bad mesh-inside arguments (triangles %lld, points %lld)", (long long)nt, (long long)np);
        return HFM_ERR_ARG;
    }
    if (np == 0) return HFM_OK;
    RC(check_finite(tri, 9 * nt, "triangles"));
    RC(check_finite(points, 3 * np, "points"));
    const int ndev = hfm_device_count();
    if (device < 0 || device >= ndev) {
        set_err("device %d not available (%d HIP devices)", device, ndev);
        return HFM_ERR_ARG;
    }
    int prev = 0;
    HIP_OK(hipGetDevice(&prev));
    HIP_OK(hipSetDevice(device));
    float *d_tri = nullptr, *d_pts = nullptr, *d_wn = nullptr;
    int* d_flags = nullptr;
    int rc = HFM_OK;
    auto fail = [&](hipError_t e, const char* what) {
        set_err("%s failed: %s", what, hipGetErrorString(e));
        rc = e == hipErrorOutOfMemory ? HFM_ERR_OOM : HFM_ERR_HIP;
    };
    hipError_t e = hipSuccess;
    if (nt > 0 && (e = hipMalloc(&d_tri, 9 * nt * sizeof(float))) != hipSuccess) fail(e, "hipMalloc");
    if (!rc && (e = hipMalloc(&d_pts, 3 * np * sizeof(float))) != hipSuccess) fail(e, "hipMalloc");
    if (!rc && (e = hipMalloc(&d_flags, np * sizeof(int))) != hipSuccess) fail(e, "hipMalloc");
    if (!rc && winding && (e = hipMalloc(&d_wn, np * sizeof(float))) != hipSuccess) fail(e, "hipMalloc");
    if (!rc && nt > 0 && (e = hipMemcpy(d_tri, tri, 9 * nt * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess)
        fail(e, "hipMemcpy");
    if (!rc && (e = hipMemcpy(d_pts, points, 3 * np * sizeof(float), hipMemcpyHostToDevice)) != hipSuccess)
        fail(e, "hipMemcpy");
    if (!rc) {
        launch_winding(d_tri, nt, d_pts, np, threshold, d_flags, d_wn, nullptr);
        if ((e = hipGetLastError()) != hipSuccess) fail(e, "kernel launch");
    }
    if (!rc && (e = hipMemcpy(flags, d_flags, np * sizeof(int), hipMemcpyDeviceToHost)) != hipSuccess)
        fail(e, "hipMemcpy");
    if (!rc && winding && (e = hipMemcpy(winding, d_wn, np * sizeof(float), hipMemcpyDeviceToHost)) != hipSuccess)
        fail(e, "hipMemcpy");
    for (void* q : {(void*)d_tri, (void*)d_pts, (void*)d_flags, (void*)d_wn})
        if (q) (void)hipFree(q);
    (void)hipSetDevice(prev);   // the caller's current device is left as it was
    return rc;
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
    if (prm->precision != HFM_PREC_FP32 && prm->precision != HFM_PREC_MIXED) {
        set_err("precision mode %d is not available", prm->precision);
        return HFM_ERR_ARG;
    }
    if (prm->dir_planes != 0 && prm->dir_planes != -1 && prm->dir_planes != 4 && prm->dir_planes != 8 &&
        prm->dir_planes != 16) {
        set_err("dir_planes %d: use 0, -1, 4, 8 or 16", prm->dir_planes);
        return HFM_ERR_ARG;
    }
    if ((prm->flags & HFM_FLAG_MG_PCG) && (prm->flags & HFM_FLAG_FUSED_CG)) {
        set_err("HFM_FLAG_MG_PCG is available with the two-pass PCG only");
        return HFM_ERR_ARG;
    }
    if (prm->flags & HFM_FLAG_MG_PCG) {
        int m = prm->n, levels = 1;
        for (; m > 8 && m % 2 == 0; ++levels) m /= 2;
        if (m > 8 || prm->n < 16 || levels > kMgMaxLevels) {
            set_err("HFM_FLAG_MG_PCG needs n = m 2^k with m <= 8, k <= %d and n >= 16 (n = %d)", kMgMaxLevels - 1,
                    prm->n);
            return HFM_ERR_ARG;
        }
    }
    if (prm->precision == HFM_PREC_MIXED && (prm->flags & HFM_FLAG_FUSED_CG)) {
        set_err("HFM_FLAG_FUSED_CG is not available with HFM_PREC_MIXED");
        return HFM_ERR_ARG;
    }
    if (prm->precision == HFM_PREC_MIXED && (prm->flags & HFM_FLAG_MACCORMACK)) {
        set_err("HFM_FLAG_MACCORMACK is available with HFM_PREC_FP32 only");
        return HFM_ERR_ARG;
    }
    const int ndev = hfm_device_count();
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
    s->n = prm->n;
    s->N = (long)prm->n * prm->n * prm->n;
    s->F = (long)(prm->n + 1) * prm->n * prm->n;
    s->heat_pred = 2;
    s->pres_pred = 8;
    s->mg_coef_dirty = 1;
    {
        // minimum tiles per PCG workgroup (the workgroup cap may raise it); HFM_PCG_CPT overrides
        const char* e = getenv("HFM_PCG_CPT");
        const int c = e ? atoi(e) : 1;
        s->pcg_cpt = c >= 1 && c <= 64 ? c : 1;
        e = getenv("HFM_PCG_FLAT_DIR");
        s->flat_dir = e && atoi(e) != 0;
        e = getenv("HFM_DIR_ICHUNK");
        const int ic = e ? atoi(e) : 0;
        s->dir_ichunk = ic >= 1 && ic <= 1024 ? ic : 0;
        // the layout asked for in the parameters (an autotuned choice) takes precedence
        if (prm->dir_planes == 4 || prm->dir_planes == 8 || prm->dir_planes == 16) {
            s->dir_ichunk = prm->dir_planes;
            s->flat_dir = 0;
        } else if (prm->dir_planes == -1) {
            s->flat_dir = 1;
        }
    }
    SceneDev& d = s->dev;
    d.mean_tail = -1;
    int rc = HFM_OK;
    hipError_t e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        set_err("hipStreamCreate: %s", hipGetErrorString(e));
        delete o;
        return HFM_ERR_HIP;
    }
    const long N = s->N, F = s->F;
#define A(ptr, cnt)                        \
    if ((rc = dalloc(o, &(ptr), (cnt)))) { \
        hfm_scene_destroy((hfm_scene*)o);  \
        return rc;                         \
    }
    for (int a = 0; a < 3; ++a) {
        A(d.u[a], F);
        A(d.us[a], F);
        A(d.q[a], F);
        A(d.kf[a], F);
        A(d.mf[a], F);
        A(d.st_k[a], F);
        A(d.st_m[a], F);
    }
    A(d.p, N);
    A(d.rhs, N);
    A(d.pdiag, N);
    A(d.alpha_s, N);
    A(d.st_alpha, N);
    A(d.T, N);
    A(d.T2, N);
    A(d.Tw, N);
    A(d.kap, N);
    A(d.alpha_p, N);
    A(d.tsum, N);
    A(d.tfix, N);
    A(d.hrhs, N);
    A(d.hdiag, N);
    A(d.flags, N);
    A(d.cg_r, 2 * N);
    A(d.cg_z, N);
    A(d.cg_d, 2 * N);
    A(d.cg_Ap, 2 * N);   // two-pass: A p; one-launch form: w = A D r, two parity buffers
    A(d.partials, std::max((long)kMaxDirBlocks + 5 * kMaxPcgBlocks, kPartialsDoubles));
    A(d.src_part, nblocks(N));
    if (prm->flags & HFM_FLAG_MG_PCG) {   // multigrid levels: halve n down to at most 8 cells per side
        d.cg_mg = 1;
        int L = 1;
        d.mg_n[0] = prm->n;
        d.mg_off[0] = 0;
        long coarse = 0;
        while (d.mg_n[L - 1] > 8 && d.mg_n[L - 1] % 2 == 0 && L < kMgMaxLevels) {
            d.mg_n[L] = d.mg_n[L - 1] / 2;
            d.mg_off[L] = coarse;
            coarse += (long)d.mg_n[L] * d.mg_n[L] * d.mg_n[L];
            ++L;
        }
        d.mg_levels = L;
        d.mg_tail = 1;
        while (d.mg_n[d.mg_tail] > kMgTailMaxN) ++d.mg_tail;   // the last level has at most 8
        A(d.mg_ax, coarse);
        A(d.mg_ay, coarse);
        A(d.mg_az, coarse);
        A(d.mg_w, coarse);
        A(d.mg_s, coarse);
        A(d.mg_e, coarse);
    }
    if (prm->flags & HFM_FLAG_MACCORMACK) {   // the forward semi-Lagrangian fields of the MacCormack step
        d.maccormack = 1;
        A(d.mc_T, N);
        for (int a = 0; a < 3; ++a) A(d.mc_u[a], F);
    }
    if (prm->precision == HFM_PREC_MIXED) {
        d.mixed = 1;
        A(d.ir_r, N);
        A(d.ir_z, N);
        A(d.ir_d, 2 * N);
        A(d.ir_q, N);
    }
    A(d.ctl, 2);
    d.nchunks = (int)((N + 8191) / 8192);
    A(d.chunk_sums, d.nchunks);
    A(s->scratch_field, F);
    A(s->umax, 4);
    A(s->bin_cnt, N + 1);
    A(d.bin_start, N + 1);
    A(d.csr_start, N + 1);
    A(s->scan_tmp, (N + 1) / 2048 + 16);
    // pairwise-sum programs for the numpy mean
    append_program(o->prog, (int)(N >= 8192 ? 8192 : N));
    if (N % 8192 != 0 && N > 8192) {
        d.mean_tail = (int)o->prog.size();
        append_program(o->prog, (int)(N % 8192));
    }
    int* prog_dev = nullptr;
    A(prog_dev, (long)o->prog.size());
#undef A
    d.mean_prog = prog_dev;
    const bool want_fused = (prm->flags & HFM_FLAG_FUSED_CG) != 0;
    const bool want_persist = want_fused && (prm->flags & HFM_FLAG_PERSISTENT_CG) != 0;
    pcg_geometry(s, want_fused, want_persist, &d.cpt, &d.nb_upd, &d.nb_dir, &d.tiles_j, &d.tiles_k, &d.dtile);
    // the one-launch iteration runs on the tiled blocks (a flat layout keeps the two-pass form); its
    // persistent form where every tile's workgroup stays resident at once
    d.cg_fused = want_fused && d.dtile > 0;
    d.cg_persist = d.cg_fused && want_persist && d.nb_dir <= kPersistMaxBlocks &&
                   d.nb_dir <= pcg_persist_capacity(d.dtile, device);
    e = hipMemcpy(prog_dev, o->prog.data(), o->prog.size() * sizeof(int), hipMemcpyHostToDevice);
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
    ctl[0].ir_maxiter = prm->heat_max_iter;
    ctl[1].tol = prm->cg_tol;
    ctl[1].maxeff = prm->cg_max_iter + (prm->cg_max_iter & 1);
    ctl[1].ir_maxiter = prm->cg_max_iter;
    e = hipMemcpy(d.ctl, ctl, sizeof(ctl), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipHostMalloc((void**)&s->mirror, 2 * sizeof(CgMirror), hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) {
        memset(s->mirror, 0, 2 * sizeof(CgMirror));
        e = hipHostGetDevicePointer((void**)&d.mirror, s->mirror, 0);
    }
    if (e != hipSuccess) {
        set_err("control block setup: %s", hipGetErrorString(e));
        hfm_scene_destroy((hfm_scene*)o);
        return HFM_ERR_HIP;
    }
    // heat.py:182-183: T, T2, Tw start at the room temperature
    std::vector<float> tamb(N, (float)prm->t_amb);
    e = hipMemcpy(d.T, tamb.data(), N * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d.T2, tamb.data(), N * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(d.Tw, tamb.data(), N * sizeof(float), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        set_err("initial temperature: %s", hipGetErrorString(e));
        hfm_scene_destroy((hfm_scene*)o);
        return HFM_ERR_HIP;
    }
    refresh_dev(s);
    s->solo = new Batch();
    s->solo->scenes.push_back(s);
    rc = batch_init(s->solo, device, s->stream, false);
    if (!rc && (prm->flags & HFM_FLAG_TIME_KERNELS)) rc = enable_timing(s->solo);
    if (rc) {
        hfm_scene_destroy((hfm_scene*)o);
        return rc;
    }
    s->current = s->solo;
    *out = (hfm_scene*)o;
    return HFM_OK;
}

int hfm_scene_destroy(hfm_scene* h) {
    if (!h) return HFM_OK;
    Owner* o = owner(h);
    Scene* s = &o->s;
    if (s->members > 0) {
        set_err("scene is held by %d batch(es); destroy them first", s->members);
        return HFM_ERR_STATE;
    }
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->solo) {
        batch_release(s->solo);
        delete s->solo;
        s->solo = nullptr;
    }
    for (void* p : o->dev)
        if (p) (void)hipFree(p);
    if (s->mirror) (void)hipHostFree(s->mirror);
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
    const long N = s->N, F = s->F;
    const float* k[3] = {kx, ky, kz};
    const float* m[3] = {mx, my, mz};
    int rc = check_finite(alpha, N, "alpha");
    for (int a = 0; a < 3 && !rc; ++a) {
        rc = check_finite(k[a], F, "k");
        if (!rc) rc = check_finite(m[a], F, "m");
    }
    if (rc) return rc;
    HIP_OK(hipStreamSynchronize(s->stream));
    HIP_OK(hipMemcpy(s->dev.st_alpha, alpha, N * sizeof(float), hipMemcpyHostToDevice));
    for (int a = 0; a < 3; ++a) {
        HIP_OK(hipMemcpy(s->dev.st_k[a], k[a], F * sizeof(float), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(s->dev.st_m[a], m[a], F * sizeof(float), hipMemcpyHostToDevice));
    }
    s->has_static = 1;
    s->mg_coef_dirty = 1;
    if (s->dev.P == 0) {
        RC(reset_particle_fields(s));
        HIP_OK(hipStreamSynchronize(s->stream));
    } else {
        s->bins_dirty = 1;  // the next gather adds the new static part to every lattice
    }
    return HFM_OK;
}

int hfm_set_fixed(hfm_scene* h, const int32_t* fixed, const float* tfix) {
    SCENE_OR_FAIL(h);
    if (!fixed || !tfix) {
        set_err("null array");
        return HFM_ERR_ARG;
    }
    const long N = s->N;
    RC(check_finite(tfix, N, "tfix"));
    std::vector<uint8_t> fl(N);
    HIP_OK(hipStreamSynchronize(s->stream));
    HIP_OK(hipMemcpy(fl.data(), s->dev.flags, N, hipMemcpyDeviceToHost));
    for (long t = 0; t < N; ++t) fl[t] = (uint8_t)((fl[t] & ~FLAG_FIXED) | (fixed[t] == 1 ? FLAG_FIXED : 0));
    HIP_OK(hipMemcpy(s->dev.flags, fl.data(), N, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(s->dev.tfix, tfix, N * sizeof(float), hipMemcpyHostToDevice));
    launch_apply_fixed(s, s->stream, s->dev.T);  // heat.py:192
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
    const long N = s->N;
    std::vector<uint8_t> fl(N);
    HIP_OK(hipStreamSynchronize(s->stream));
    HIP_OK(hipMemcpy(fl.data(), s->dev.flags, N, hipMemcpyDeviceToHost));
    for (long t = 0; t < N; ++t) fl[t] = (uint8_t)((fl[t] & ~FLAG_SOURCE) | (mask[t] == 1 ? FLAG_SOURCE : 0));
    HIP_OK(hipMemcpy(s->dev.flags, fl.data(), N, hipMemcpyHostToDevice));
    s->rate = rate;
    s->t_cap = t_cap;
    refresh_dev(s);
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
    SceneDev& d = s->dev;
    if (P > s->P_cap) {
        void* olds[] = {d.pos, d.vel, d.temp, d.melt, d.yield, d.visc, s->pkey, s->pslot, d.sorted, d.qpos,
                        d.csr_p, d.csr_w};
        for (void* q : olds)
            if (q) dfree(o, q);
        d.pos = d.vel = d.temp = d.melt = d.yield = d.visc = d.qpos = d.csr_w = nullptr;
        s->pkey = s->pslot = d.sorted = d.csr_p = nullptr;
        s->P_cap = 0;
        if ((rc = dalloc(o, &d.pos, 3 * P)) || (rc = dalloc(o, &d.vel, 3 * P)) || (rc = dalloc(o, &d.temp, P)) ||
            (rc = dalloc(o, &d.melt, P)) || (rc = dalloc(o, &d.yield, P)) || (rc = dalloc(o, &d.visc, P)) ||
            (rc = dalloc(o, &s->pkey, P)) || (rc = dalloc(o, &s->pslot, P)) || (rc = dalloc(o, &d.sorted, P)) ||
            (rc = dalloc(o, &d.qpos, 3 * P)) || (rc = dalloc(o, &d.csr_p, P <= kCsrMaxParticles ? 8 * P : 1)) ||
            (rc = dalloc(o, &d.csr_w, P <= kCsrMaxParticles ? 8 * P : 1))) {
            s->csr_cap = 0;
            d.P = 0;
            refresh_dev(s);
            return rc;
        }
        s->P_cap = P;
        s->csr_cap = P <= kCsrMaxParticles ? 8 * P : 1;
    }
    d.P = (int)P;
    refresh_dev(s);
    if (P > 0) {
        HIP_OK(hipMemcpy(d.pos, pos, 3 * P * sizeof(float), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(d.vel, vel, 3 * P * sizeof(float), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(d.temp, temp, P * sizeof(float), hipMemcpyHostToDevice));
        std::vector<float> init(P, (float)s->prm.yield_solid);
        HIP_OK(hipMemcpy(d.yield, init.data(), P * sizeof(float), hipMemcpyHostToDevice));  // mpm.py:42
        std::fill(init.begin(), init.end(), (float)s->prm.mu_solid);
        HIP_OK(hipMemcpy(d.visc, init.data(), P * sizeof(float), hipMemcpyHostToDevice));   // mpm.py:43
        HIP_OK(hipMemset(d.melt, 0, P * sizeof(float)));                                     // solver.py:61
        s->bins_dirty = 1;
    } else {
        RC(reset_particle_fields(s));
        HIP_OK(hipStreamSynchronize(s->stream));
    }
    return HFM_OK;
}

int hfm_update_particles(hfm_scene* h, const float* pos, const float* vel) {
    SCENE_OR_FAIL(h);
    const long P = s->dev.P;
    if (P == 0) return HFM_OK;
    if (!pos || !vel) {
        set_err("null array");
        return HFM_ERR_ARG;
    }
    RC(check_finite(pos, 3 * P, "pos"));
    RC(check_finite(vel, 3 * P, "vel"));
    HIP_OK(hipMemcpyAsync(s->dev.pos, pos, 3 * P * sizeof(float), hipMemcpyHostToDevice, s->stream));
    HIP_OK(hipMemcpyAsync(s->dev.vel, vel, 3 * P * sizeof(float), hipMemcpyHostToDevice, s->stream));
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
    return batch_substep(s->solo, dt);
}

int hfm_run(hfm_scene* h, double dt, int32_t K, double* ms) {
    SCENE_OR_FAIL(h);
    if (K < 0 || !(dt > 0) || !(dt < 1e3)) {
        set_err("bad run arguments");
        return HFM_ERR_ARG;
    }
    return batch_run(s->solo, dt, K, ms);
}

int64_t hfm_field_size(const hfm_scene* h, int32_t field) {
    if (!h) return -1;
    Scene* s = &owner(const_cast<hfm_scene*>(h))->s;
    if (field >= HFM_FIELD_AX && field <= HFM_FIELD_AZ) return s->F;
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
        launch_face_fraction(s, s->stream, field - HFM_FIELD_AX, s->scratch_field);
        CHECK_LAUNCH();
        HIP_OK(hipStreamSynchronize(s->stream));
        HIP_OK(hipMemcpy(out, s->scratch_field, s->F * sizeof(float), hipMemcpyDeviceToHost));
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
    RC(check_finite(in, size, "field"));
    HIP_OK(hipStreamSynchronize(s->stream));
    HIP_OK(hipMemcpy(p, in, size * sizeof(float), hipMemcpyHostToDevice));
    if (field == HFM_FIELD_ALPHA_S) s->mg_coef_dirty = 1;
    return HFM_OK;
}

int hfm_get_cell_velocity(hfm_scene* h, float* out) {
    SCENE_OR_FAIL(h);
    if (!out) {
        set_err("null array");
        return HFM_ERR_ARG;
    }
    float* tmp = nullptr;
    HIP_OK(hipMalloc(&tmp, 3 * s->N * sizeof(float)));
    launch_cell_velocity(s, s->stream, tmp);
    hipError_t e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) e = hipMemcpy(out, tmp, 3 * s->N * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(tmp);
    HIP_OK(e);
    return HFM_OK;
}

int hfm_get_particles(hfm_scene* h, float* temp, float* melt, float* yield_stress, float* viscosity) {
    SCENE_OR_FAIL(h);
    HIP_OK(hipStreamSynchronize(s->stream));
    if (s->dev.P == 0) return HFM_OK;
    const size_t bytes = (size_t)s->dev.P * sizeof(float);
    if (temp) HIP_OK(hipMemcpy(temp, s->dev.temp, bytes, hipMemcpyDeviceToHost));
    if (melt) HIP_OK(hipMemcpy(melt, s->dev.melt, bytes, hipMemcpyDeviceToHost));
    if (yield_stress) HIP_OK(hipMemcpy(yield_stress, s->dev.yield, bytes, hipMemcpyDeviceToHost));
    if (viscosity) HIP_OK(hipMemcpy(viscosity, s->dev.visc, bytes, hipMemcpyDeviceToHost));
    return HFM_OK;
}

int hfm_get_particle_bins(hfm_scene* h, int32_t* bin, int32_t* base, float* frac) {
    SCENE_OR_FAIL(h);
    if (!bin || !base || !frac) {
        set_err("null array");
        return HFM_ERR_ARG;
    }
    const long P = s->dev.P;
    if (P == 0) return HFM_OK;
    if (s->bins_dirty) {   // bin now, exactly as the next substep would (its state kept in step)
        launch_bin_particles(s, s->stream);
        s->csr_state = 0;
        s->mg_coef_dirty = 1;
    }
    int *d_bin = nullptr, *d_base = nullptr;
    float* d_frac = nullptr;
    hipError_t e = hipMalloc(&d_bin, P * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&d_base, 12 * P * sizeof(int));
    if (e == hipSuccess) e = hipMalloc(&d_frac, 12 * P * sizeof(float));
    if (e == hipSuccess) {
        launch_particle_bins(s, s->stream, d_bin, d_base, d_frac);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    if (e == hipSuccess) e = hipMemcpy(bin, d_bin, P * sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(base, d_base, 12 * P * sizeof(int), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(frac, d_frac, 12 * P * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(d_bin);
    (void)hipFree(d_base);
    (void)hipFree(d_frac);
    HIP_OK(e);
    return HFM_OK;
}

int hfm_diagnostics(hfm_scene* h, double* divergence_max, double* speed_max) {
    SCENE_OR_FAIL(h);
    HIP_OK(hipMemsetAsync(s->umax, 0, 4 * sizeof(unsigned), s->stream));
    launch_divergence_scale1(s, s->stream, s->scratch_field);
    launch_abs_max(s->stream, s->scratch_field, s->N, &s->umax[0]);
    for (int a = 0; a < 3; ++a) launch_abs_max(s->stream, s->dev.u[a], s->F, &s->umax[1]);
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
    return kernel_time(s->solo, cls, total_ms, launches);
}

int hfm_kernel_time_reset(hfm_scene* h) {
    SCENE_OR_FAIL(h);
    return kernel_time_reset(s->solo);
}

int hfm_set_kernel_timing(hfm_scene* h, int32_t on) {
    SCENE_OR_FAIL(h);
    return set_timing(s->solo, on);
}

int64_t hfm_scene_bytes(const hfm_scene* h) {
    if (!h) return -1;
    const Owner* o = reinterpret_cast<const Owner*>(h);
    int64_t total = 0;
    for (size_t b : o->dev_bytes) total += (int64_t)b;
    return total;
}

int hfm_synchronize(hfm_scene* h) {
    SCENE_OR_FAIL(h);
    HIP_OK(hipStreamSynchronize(s->stream));
    return HFM_OK;
}

// ---- batches -------------------------------------------------------------------------------

int hfm_batch_create(hfm_scene* const* scenes, int32_t count, hfm_batch** out) {
    if (!scenes || !out || count < 1 || count > kMaxBatch) {
        set_err("bad batch arguments (count=%d, at most %d)", count, kMaxBatch);
        return HFM_ERR_ARG;
    }
    *out = nullptr;
    std::vector<Scene*> list;
    for (int i = 0; i < count; ++i) {
        if (!scenes[i]) {
            set_err("null scene at %d", i);
            return HFM_ERR_ARG;
        }
        Scene* s = &owner(scenes[i])->s;
        if (std::find(list.begin(), list.end(), s) != list.end()) {
            set_err("scene %d appears twice in the batch", i);
            return HFM_ERR_ARG;
        }
        if (!list.empty() && s->device != list[0]->device) {
            set_err("scenes of a batch must be on one device");
            return HFM_ERR_ARG;
        }
        if (!list.empty() && s->prm.precision != list[0]->prm.precision) {
            set_err("scenes of a batch must have one precision mode");
            return HFM_ERR_ARG;
        }
        list.push_back(s);
    }
    HIP_OK(hipSetDevice(list[0]->device));
    hfm_batch* hb = new hfm_batch();
    Batch* b = &hb->b;
    b->user = true;
    b->scenes = list;
    hipStream_t st = nullptr;
    hipError_t e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (e != hipSuccess) {
        set_err("hipStreamCreate: %s", hipGetErrorString(e));
        delete hb;
        return HFM_ERR_HIP;
    }
    int rc = batch_init(b, list[0]->device, st, true);
    if (rc) {
        batch_release(b);
        delete hb;
        return rc;
    }
    for (Scene* s : list) s->members += 1;
    *out = hb;
    return HFM_OK;
}

int hfm_batch_set_scenes(hfm_batch* h, hfm_scene* const* scenes, int32_t count) {
    BATCH_OR_FAIL(h);
    if ((count > 0 && !scenes) || count < 0 || count > kMaxBatch) {
        set_err("bad batch arguments (count=%d, at most %d)", count, kMaxBatch);
        return HFM_ERR_ARG;
    }
    std::vector<Scene*> list;
    for (int i = 0; i < count; ++i) {
        if (!scenes[i]) {
            set_err("null scene at %d", i);
            return HFM_ERR_ARG;
        }
        Scene* s = &owner(scenes[i])->s;
        if (std::find(list.begin(), list.end(), s) != list.end()) {
            set_err("scene %d appears twice in the batch", i);
            return HFM_ERR_ARG;
        }
        if (s->device != b->device) {
            set_err("scenes of a batch must be on one device");
            return HFM_ERR_ARG;
        }
        if (!list.empty() && s->prm.precision != list[0]->prm.precision) {
            set_err("scenes of a batch must have one precision mode");
            return HFM_ERR_ARG;
        }
        list.push_back(s);
    }
    HIP_OK(hipStreamSynchronize(b->stream));
    for (Scene* s : b->scenes) {
        if (s->current == b) s->current = s->solo;
        s->members -= 1;
    }
    for (Scene* s : list) s->members += 1;
    b->scenes = list;
    b->versions.clear();  // the next substep rebuilds the table and the maps (in the same buffers)
    return HFM_OK;
}

int hfm_batch_destroy(hfm_batch* h) {
    if (!h) return HFM_OK;
    Batch* b = &h->b;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (Scene* s : b->scenes) {
        if (s->current == b) s->current = s->solo;
        s->members -= 1;
    }
    batch_release(b);
    delete h;
    return HFM_OK;
}

int hfm_batch_substep(hfm_batch* h, double dt) {
    BATCH_OR_FAIL(h);
    if (b->scenes.empty()) return HFM_OK;
    if (!(dt > 0) || !(dt < 1e3)) {
        set_err("bad dt %g", dt);
        return HFM_ERR_ARG;
    }
    return batch_substep(b, dt);
}

int hfm_batch_run(hfm_batch* h, double dt, int32_t K, double* ms) {
    BATCH_OR_FAIL(h);
    if (b->scenes.empty()) {
        if (ms) *ms = 0.0;
        return HFM_OK;
    }
    if (K < 0 || !(dt > 0) || !(dt < 1e3)) {
        set_err("bad run arguments");
        return HFM_ERR_ARG;
    }
    return batch_run(b, dt, K, ms);
}

int hfm_batch_synchronize(hfm_batch* h) {
    BATCH_OR_FAIL(h);
    HIP_OK(hipStreamSynchronize(b->stream));
    return HFM_OK;
}

int hfm_batch_set_kernel_timing(hfm_batch* h, int32_t on) {
    BATCH_OR_FAIL(h);
    return set_timing(b, on);
}

int hfm_batch_kernel_time(hfm_batch* h, int32_t cls, double* total_ms, int64_t* launches) {
    BATCH_OR_FAIL(h);
    return kernel_time(b, cls, total_ms, launches);
}

int hfm_batch_kernel_time_reset(hfm_batch* h) {
    BATCH_OR_FAIL(h);
    return kernel_time_reset(b);
}

}  // extern "C"
