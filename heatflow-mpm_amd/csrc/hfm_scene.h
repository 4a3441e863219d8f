// hfm_scene.h — scenes, batches and kernel launchers of libhfm.so.
//
// Every substep kernel runs over a batch: a set of scenes (grids of any size on one device) that
// advance together.  Each launch covers all scenes of the batch: workgroup b reads a 64-bit
// entry of a block map (scene, sub-range, local workgroup) and the scene's SceneDev descriptor,
// so one launch per stage serves every scene and the fixed cost of a launch is shared.  A
// single scene runs as a batch of one.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <hip/hip_ext.h>

#include <vector>

#include "../../include/hfm.h"

namespace hfm {

// Device-resident control block of one PCG solve (warp.optim.linear.cg semantics).  Written by
// local workgroup 0 of a scene in a kernel and read by later kernels only.
struct CgCtl {
    float slot_rr[2];  // ||r||^2 after iteration k, in slot k & 1
    float slot_rz[2];  // r.z after iteration k, in slot k & 1
    double slot_pap[2];  // fused form: p.Ap of iteration k, in slot k & 1
    float atol_sq;     // fp32 threshold on ||r||^2
    int it;            // iterations done when the solve stopped
    int done;          // 1: converged or iteration cap reached; every later CG kernel skips the scene
    int converged;     // 1 converged, 0 cap reached, -1 stopped on a breakdown (see hfm_stats)
    int brk;           // set by pass B when p.Ap is not positive or alpha not finite; the next pass A
                       // stops the solve there
    int maxeff;        // iteration cap rounded up to the 2-iteration cycle
    int geom_dirty;    // ctl[0] only: particles were re-binned, the next gather rebuilds every lattice
    int csr_ready;     // ctl[0] only: the cell lattice's contribution list matches the binned particles
    double tol;        // relative tolerance
    double atol;       // max(tol * ||b||, tol)
    float mean;        // pressure rhs mean to remove (numpy float32 mean)
    unsigned ticket2;  // arrival counter of the rhs-mean chunks
    // mixed precision (iterative refinement): written by the scale kernel of each refinement step
    float ir_gl;       // (atol / sigma)^2: the global threshold in the inner solve's units
    float ir_step;     // sigma * c: x += ir_step * alpha * p of the inner solve
    int ir_total;      // inner iterations of this solve so far (written when an inner solve stops)
    int ir_outer;      // refinement steps of this solve so far
    int ir_end;        // 1: the refinement has converged or reached the cap
    int ir_inner_conv; // 1: the last inner solve reached its tolerance (written by the residual launch)
    int ir_maxiter;    // iteration cap (inner iterations over the whole solve)
    double ir_rr2[2];  // ||r||^2 (true residual) of refinement step s, in slot s & 1
    double src_dt;     // ctl[0]: sum of the source step's temperature rises of this substep (fp64)
    // multigrid-preconditioned fp32 solves: smallest ||r||^2 so far and its iteration, after iteration k in
    // slot k & 1 (the floor stop)
    float slot_rrmin[2];
    int slot_kmin[2];
};

// Pinned host memory written by the device with system-scope stores (results of a solve),
// polled by the host.
struct CgMirror {
    int it;
    int done;
    int converged;
    int outer;         // mixed precision: refinement steps
    double rr;
    double atol;
    double mean;
    int end;           // mixed precision: 1 when `done` ends the whole solve, 0 when one inner solve
    int pad;
    double src_dt;     // [0]: CgCtl::src_dt of this substep (the heater energy balance)
};

// Pointer members of the per-scene descriptor.  The kernels read the descriptor from device
// memory, so the compiler cannot see where its pointers point; in the device pass the kernels'
// view of it types them as global memory, which makes every access a global (not flat)
// load/store with a scalar base.  Host code and the host pass see plain pointers.
#if defined(__HIP_DEVICE_COMPILE__)
#define HFM_GLOBAL __attribute__((address_space(1)))
#else
#define HFM_GLOBAL
#endif
struct HostPtrs {
    template <class T>
    using p = T*;
};
struct DevPtrs {
    template <class T>
    using p = HFM_GLOBAL T*;
};

// Per-scene view read by the kernels (one entry per scene of a batch, in device memory).
// Parameters are fp32 as the reference passes them to its kernels; rho is also kept in double
// because rho/dt and dt/rho are formed in double on the reference's host (grid.py:237,251).
template <class S>
struct SceneView {
    template <class T>
    using ptr = typename S::template p<T>;
    int n, P;
    long N, F;
    float h, hh;
    float rhh;               // RN(1 / hh)
    int fastdiv;             // x / hh through div_const's corrected product (checked on the host)
    float rh;                // RN(1 / h)
    int fast_h;              // pos / h through div_const's corrected product (checked on the host)
    float rho, t_amb, g, amin, ka, km, frac, k0, t_sol, t_liq, ys, mus, mul, t_max, rate, t_cap;
    int rate_on, buoyancy;
    double rho_d;
    ptr<float> u[3], us[3], q[3], kf[3], mf[3], st_k[3], st_m[3];
    ptr<float> p, rhs, pdiag, alpha_s, st_alpha;
    ptr<float> T, T2, Tw, kap, alpha_p, tsum, tfix, hrhs, hdiag;
    ptr<uint8_t> flags;
    ptr<float> cg_r, cg_z, cg_d, cg_Ap;
    ptr<double> partials;    // two-pass: [kMaxDirBlocks] pass A, then [3][nb_upd] pass B / init;
                             // fused: two parity buffers of [kFusedRows][kMaxDirBlocks]
    ptr<CgCtl> ctl;          // [0] heat, [1] pressure
    ptr<double> src_part;    // [nblocks(N)] per-workgroup sums of the source step's rises (heat prep)
    ptr<CgMirror> mirror;    // device view of the pinned mirrors, [0] heat, [1] pressure
    int cpt, nb_upd, nb_dir, tiles_j, tiles_k, dtile;
    int cg_fused;            // pressure PCG in one launch per iteration (tiled scenes only)
    int cg_persist;          // ... and all its iterations in one cooperative launch (HFM_FLAG_PERSISTENT_CG)
    ptr<float> chunk_sums;
    ptr<const int> mean_prog;
    int mean_tail;           // offset of the tail chunk's program in mean_prog, -1 if none
    int nchunks;
    ptr<float> pos, vel;
    ptr<float> temp, melt, yield, visc;
    ptr<int> bin_start, sorted;
    ptr<float> qpos;         // pos / h per particle, in bin order (3 per particle)
    // contribution list of the cell lattice while the particles stay put: the terms w of node c
    // are csr_w[csr_start[c] .. csr_start[c + 1]) of particles csr_p[...], in the gather's order
    ptr<int> csr_start;      // [N + 1]
    ptr<int> csr_p;          // [8 P]
    ptr<float> csr_w;        // [8 P]
    int mixed;
    int maccormack;          // HFM_FLAG_MACCORMACK: limited MacCormack advection of T and u, v, w
    ptr<float> mc_T;         // [N] forward semi-Lagrangian T of this substep (MacCormack only)
    ptr<float> mc_u[3];      // [F] forward semi-Lagrangian u, v, w of this substep (MacCormack only)               // HFM_PREC_MIXED
    int cg_mg;               // pressure PCG preconditioned by a multigrid V-cycle (HFM_FLAG_MG_PCG)
    int mg_levels;           // grid levels, level 0 the scene's grid, each next one halved
    int mg_tail;             // first level of at most kMgTailMaxN cells per side (solved in LDS)
    int mg_n[8];             // cells per side of each level
    long mg_off[8];          // first cell of level l >= 1 in the packed coarse-level arrays
    ptr<float> mg_ax, mg_ay, mg_az, mg_w;   // coarse levels: -x, -y, -z face coefficients, omega / sum
    ptr<float> mg_s, mg_e;   // coarse levels: right-hand side and correction of the V-cycle
    ptr<_Float16> ir_r, ir_z, ir_d, ir_q;  // inner solve: r, z, p (two parity buffers), A p
};
using SceneDev = SceneView<HostPtrs>;  // host copy, written by the host
using SceneG = SceneView<DevPtrs>;     // the kernels' view of the same bytes

// Block-map entries: scene (24 bits) | sub-range (8 bits) | local workgroup (32 bits).
typedef unsigned long long MapEntry;
constexpr MapEntry kNoBlock = ~0ull;
constexpr int kMaxBatch = 4096;
__host__ __device__ inline MapEntry map_entry(int s, int sub, long l) {
    return ((MapEntry)s << 40) | ((MapEntry)sub << 32) | (MapEntry)(unsigned)l;
}

// Where a launch finds its scenes.  A batch passes a block map and the descriptor table: workgroup
// b reads map[b] (scene, sub-range, local workgroup), then the scene's descriptor.  A batch of one
// passes the descriptor itself as a kernel argument (sub-range = blockIdx.y, local workgroup =
// blockIdx.x), so no dependent loads stand between the launch and the first data load.
// The table and the map are read-only for the whole launch; in the device pass they are typed as
// constant memory, so their uniform reads are scalar loads into SGPRs.
#if defined(__HIP_DEVICE_COMPILE__)
#define HFM_CONST __attribute__((address_space(4)))
#else
#define HFM_CONST
#endif
struct BatchSrc {
    const HFM_CONST SceneG* tab;
    const HFM_CONST MapEntry* map;
};
struct SoloSrc {
    SceneG s;
};

__device__ __forceinline__ bool resolve(const BatchSrc& src, int& sub, int& lb, const HFM_CONST SceneG*& a) {
    const MapEntry e = src.map[blockIdx.x];
    if (e == kNoBlock) return false;
    sub = (int)((e >> 32) & 0xFFu);
    lb = (int)(e & 0xFFFFFFFFu);
    a = &src.tab[e >> 40];
    return true;
}
__device__ __forceinline__ bool resolve(const SoloSrc& src, int& sub, int& lb, const SceneG*& a) {
    sub = (int)blockIdx.y;
    lb = (int)blockIdx.x;
    a = &src.s;
    return true;
}

// In a batch, scenes whose solve has stopped are skipped before any of their data is loaded:
// one uniform scalar load of the flag (a batch keeps launching until its slowest scene converges,
// so most workgroups may belong to finished scenes).  A wave that leaves early is not waited for
// at later barriers, and the workgroups that still run reach the same stop decision from the
// partial sums, so the outcome does not depend on when the flag is seen.  A batch of one needs no
// such check (its launches end with its solve) and saves the round trip.
template <class Src>
__device__ __forceinline__ bool batch_solve_done(const Src&, const CgCtl* c) {
    return false;
}
template <>
__device__ __forceinline__ bool batch_solve_done<BatchSrc>(const BatchSrc&, const CgCtl* c) {
    return c->done != 0;
}

template <class S>
struct SrcScene;
template <>
struct SrcScene<BatchSrc> {
    using type = const HFM_CONST SceneG*;
};
template <>
struct SrcScene<SoloSrc> {
    using type = const SceneG*;
};

// Workgroup prologue of a scene kernel: `a` is the scene's descriptor, `sub` the sub-range and
// `lb` the workgroup's index inside it.  Padding entries of a block map end the workgroup at once.
#define HFM_ENTRY(src)                                   \
    int sub, lb;                                         \
    typename ::hfm::SrcScene<decltype(src)>::type ap_;   \
    if (!::hfm::resolve(src, sub, lb, ap_)) return;      \
    const auto& a = *ap_;                                \
    (void)sub;                                           \
    (void)lb;

// Launch families: how many workgroups each scene needs in each sub-range.
enum Family {
    FAM_CELL = 0,   // one workgroup per 256 cells
    FAM_FACE3,      // 3 axes x one workgroup per 256 faces
    FAM_P2G,        // cell lattice + 3 face lattices (scenes with particles)
    FAM_MEAN,       // one workgroup per 8192-cell rhs chunk
    FAM_PCG,        // PCG pass B / init / flat pass A workgroups
    FAM_G2P,        // one workgroup per 256 particles (scenes with particles)
    FAM_DIR4,       // tiled pressure pass A, 16x16x4 blocks (1024 threads)
    FAM_DIR8,
    FAM_DIR16,
    FAM_DIRFLAT,    // pressure pass A without tiling (PCG geometry)
    FAM_FUSED4,     // fused pressure iteration, 16x16x4 blocks (scenes with cg_fused)
    FAM_FUSED8,
    FAM_FUSED16,
    FAM_MG0,        // multigrid level l = 0..7 cells, one workgroup per 256 (scenes with cg_mg)
    FAM_MG1,
    FAM_MG2,
    FAM_MG3,
    FAM_MG4,
    FAM_MG5,
    FAM_MG6,
    FAM_MG7,
    FAM_MGC,        // the multigrid tail (levels of at most 12^3 cells), one workgroup per scene
    FAM_COUNT
};

struct Batch;

struct Scene {
    hfm_params prm;
    int device;
    hipStream_t stream;     // the scene's own stream (its solo batch runs on it)
    Batch* solo;            // the scene as a batch of one
    Batch* current;         // the batch whose stream last ran work for the scene
    int members;            // user batches that hold the scene
    SceneDev dev;           // descriptor (host copy)
    CgMirror* mirror;       // pinned result mirrors, [0] heat, [1] pressure (host view)
    int n;
    long N, F;
    int has_static;
    double rate, t_cap;
    long P_cap;
    int* pkey;
    int* pslot;
    int* bin_cnt;
    int* scan_tmp;
    int bins_dirty;
    int csr_state;           // contribution list: 0 none, 1 built in this substep, 2 in use
    long csr_cap;            // terms the list arrays hold (8 P of the allocation, or 1)
    int mg_coef_dirty;       // the solid fraction changed since the multigrid coefficients were formed
    unsigned version;       // bumped whenever the descriptor changes
    hfm_stats stats;
    int heat_pred, pres_pred;
    float* scratch_field;   // face-sized temporary for getters and diagnostics
    unsigned* umax;
    int flat_dir, dir_ichunk, pcg_cpt;
};

struct Batch {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool user = false;                       // created through the batch API
    std::vector<Scene*> scenes;
    SceneDev* table = nullptr;               // device copy of the descriptors
    size_t table_cap = 0;                    // descriptors the table buffer holds
    MapEntry* maps[FAM_COUNT] = {};          // device block maps
    int map_len[FAM_COUNT] = {};
    size_t map_cap[FAM_COUNT] = {};          // entries each map buffer holds (buffers only grow)
    std::vector<unsigned> versions;          // scene versions the maps and table were built for
    std::vector<MapEntry> hmap[FAM_COUNT];   // host copies of the maps
    std::vector<std::pair<size_t, size_t>> ranges[FAM_COUNT];  // per scene: [begin, end) in the map
    // maps restricted to the scenes whose solve still runs, in three rotating buffers (device and
    // pinned staging), so a buffer is rewritten only after the launches that read it completed
    MapEntry* fmap[3][FAM_COUNT] = {};
    size_t fmap_cap[3][FAM_COUNT] = {};
    MapEntry* fstage[3] = {};
    size_t fstage_cap[3] = {};
    int fslot = 0;
    MapEntry* pmap = nullptr;                // persistent groups' block maps (device, pinned staging)
    size_t pmap_cap = 0;
    MapEntry* pstage = nullptr;
    size_t pstage_cap = 0;
    unsigned long long maps_epoch = 0;       // bumped whenever batch_prepare rebuilds the table and maps
    // the multigrid V-cycle captured as a HIP graph, [0] fp32, [1] the mixed form's inner solve: valid for
    // the tail depth and maps epoch it was captured with
    hipGraphExec_t mg_exec[2] = {};
    int mg_exec_tail[2] = {};
    unsigned long long mg_exec_epoch[2] = {};
    unsigned long long* pace = nullptr;      // pinned: (solve tag << 32) | iterations completed
    unsigned long long* pace_dev = nullptr;
    unsigned solve_tag = 0;
    // kernel timing (HFM_FLAG_TIME_KERNELS)
    int timing = 0;
    hipEvent_t* ev_pool = nullptr;
    int ev_cap = 0, ev_used = 0;
    int* ev_class = nullptr;
    double kt_ms[HFM_KT_COUNT] = {};
    long long kt_n[HFM_KT_COUNT] = {};
};

struct Launch {
    const SceneDev* tab;
    const MapEntry* map;
    int blocks;             // workgroups of the batched launch (0: nothing to launch)
    hipStream_t stream;
    const SceneDev* solo;   // batch of one: its descriptor, launched as a kernel argument
    dim3 sgrid;             // batch of one: (workgroups per sub-range, sub-ranges)
};

Launch family(const Batch* b, Family f);

// Launch a scene kernel over a family, as a batch (map + table) or as a batch of one (descriptor
// argument); with ev0/ev1 the launch carries its own start/stop events (hipExtLaunchKernelGGL).
template <class KB, class KS, class... Args>
inline void launch_scene_kernel(KB kbatch, KS ksolo, const Launch& L, int threads, hipEvent_t ev0, hipEvent_t ev1,
                                Args... args) {
    if (L.blocks <= 0) return;
    if (L.solo) {
        const SoloSrc src{*reinterpret_cast<const SceneG*>(L.solo)};
        if (ev0)
            hipExtLaunchKernelGGL(ksolo, L.sgrid, dim3(threads), 0, L.stream, ev0, ev1, 0, src, args...);
        else
            hipLaunchKernelGGL(ksolo, L.sgrid, dim3(threads), 0, L.stream, src, args...);
    } else {
        BatchSrc src;
        src.tab = (const HFM_CONST SceneG*)(const void*)L.tab;
        src.map = (const HFM_CONST MapEntry*)(const void*)L.map;
        if (ev0)
            hipExtLaunchKernelGGL(kbatch, dim3(L.blocks), dim3(threads), 0, L.stream, ev0, ev1, 0, src, args...);
        else
            hipLaunchKernelGGL(kbatch, dim3(L.blocks), dim3(threads), 0, L.stream, src, args...);
    }
}
#define HFM_LAUNCH(K, L, threads, ev0, ev1, ...) \
    ::hfm::launch_scene_kernel(K<::hfm::BatchSrc>, K<::hfm::SoloSrc>, L, threads, ev0, ev1, ##__VA_ARGS__)
#define HFM_LAUNCH_T(K, T, L, threads, ev0, ev1, ...) \
    ::hfm::launch_scene_kernel(K<T, ::hfm::BatchSrc>, K<T, ::hfm::SoloSrc>, L, threads, ev0, ev1, ##__VA_ARGS__)
#define HFM_LAUNCH_T2(K, T1, T2, L, threads, ev0, ev1, ...)                                                    \
    ::hfm::launch_scene_kernel(K<T1, T2, ::hfm::BatchSrc>, K<T1, T2, ::hfm::SoloSrc>, L, threads, ev0, ev1, \
                               ##__VA_ARGS__)

// ---- batched kernel launchers (hfm_fields.hip) ----
void launch_heat_prep(const Launch& L, float dt);
void launch_advect_T(const Launch& L, float dt);
void launch_g2p(const Launch& L);
void launch_air_faces(const Launch& L, float dt);
void launch_divergence(const Launch& L, double dt);
void launch_rhs_mean(const Launch& L);
void launch_project(const Launch& L, double dt);

// ---- single-scene helpers (hfm_fields.hip) ----
void launch_divergence_scale1(const Scene* s, hipStream_t st, float* out);
void launch_cell_velocity(const Scene* s, hipStream_t st, float* out);
void launch_face_fraction(const Scene* s, hipStream_t st, int axis, float* out);
void launch_abs_max(hipStream_t st, const float* a, long count, unsigned* slot);
void launch_apply_fixed(const Scene* s, hipStream_t st, float* T);

// ---- particles (hfm_p2g.hip) ----
void launch_bin_particles(Scene* s, hipStream_t st);
void launch_particle_bins(const Scene* s, hipStream_t st, int* bin, int* base, float* frac);
void launch_p2g(const Launch& L);
constexpr long kCsrMaxParticles = 1L << 27;   // scenes with more keep the general gather every substep
void launch_csr_build(Scene* s, hipStream_t st);
void launch_mc_T(const Launch& L, float dt);
void launch_mc_faces(const Launch& L, float dt);
void launch_winding(const float* tri, long nt, const float* pts, long np, float threshold, int* flags, float* wn,
                    hipStream_t st);
// the message hfm_last_error() returns on this thread; HFM_ERR_NONFINITE (and the message) if a host
// array holds a value that is not finite
void set_error(const char* fmt, ...);
int check_finite_host(const float* a, long n, const char* what);
void launch_p2g_listed(const Launch& L);

// ---- PCG (hfm_pcg.hip); op 0 = heat, 1 = pressure ----
void pcg_geometry(const Scene* s, bool fused, bool persist, int* cpt, int* nb_upd, int* nb_dir, int* tiles_j,
                  int* tiles_k, int* dtile);
void launch_pcg_init(const Launch& L, int op, float dt);
// pace: when non-null, workgroup 0 posts (tag << 32) | k to it as the launch starts
// ir: the inner iteration of a mixed-precision refinement (FP16 vectors, scaled operator)
void launch_pcg_dir(const Launch& L, Family fam, int op, float dt, int k, unsigned long long* pace, unsigned tag,
                    hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr, bool ir = false);
void launch_pcg_upd(const Launch& L, int op, int k, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr,
                    bool ir = false);
// mixed precision: true residual of a refinement step (first: also b.b and the rhs mean removal),
// then its scaling into the inner solve's FP16 vectors (or the end of the solve)
void launch_ir_residual(const Launch& L, int op, float dt, bool first);
void launch_ir_scale(const Launch& L, int op, int step);   // step: refinement step, 0 the first

// ---- multigrid preconditioner of the pressure PCG (hfm_mg.hip) ----
constexpr int kMgMaxLevels = 8;
constexpr int kMgTailMaxN = 12;   // levels of at most 12^3 cells are solved in one workgroup
void launch_mg_coefficients(const Batch* b, int levels);
// the V-cycle of every multigrid scene of the batch (z = M r and its r.z partial sums): its launches one
// by one, or with HFM_MG_GRAPH=1 one replay of the batch's captured graph (captured on first use and
// again after the batch's maps were rebuilt)
int launch_mg_vcycle(Batch* b, int tail, bool ir = false);
// fused pressure form: first-iteration scalars after launch_pcg_init, then one launch per iteration
constexpr int kFusedRows = 5;
void launch_pcg_init_fused(const Launch& L);
void launch_pcg_fused(const Launch& L, Family fam, int k, unsigned long long* pace, unsigned tag,
                      hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// persistent form: every iteration of a solo scene's pressure solve in one cooperative launch of
// nb_dir workgroups (D planes per tile), after launch_pcg_init_fused; the capacity is the most such
// workgroups the device keeps resident at once (0: no cooperative launch, or no kernel for D)
constexpr int kPersistSpinLimit = 1 << 22;   // polls of one partial-sum wait before the solve gives up
// at most one workgroup per thread of a workgroup (each thread polls one workgroup's sums)
constexpr int kPersistMaxBlocks = 1024;
// the persistent form's partial-sum slots, [3][3 rows][nb_dir] doubles of the partial-sum buffer (the
// one-launch form's second parity buffer), and the pattern of an empty entry (a signalling NaN)
constexpr long kPersistSlots = (long)kFusedRows * kMaxDirBlocks;
constexpr unsigned long long kEmptySum = 0x7FF4A5A5A5A5A5A5ull;
// Every workgroup posts its sums to kPersistReplicas copies of the slots, kPersistReplicaStride doubles
// apart (an odd number of cache lines, so the copies fall on different pages), and polls the copy of
// its XCD: the polls of the whole device spread over eight sets of pages instead of one.
constexpr int kPersistReplicas = 8;
constexpr long kPersistReplicaStride = 9L * kPersistMaxBlocks + 16 * 37;
constexpr long kPartialsDoubles = kPersistSlots + kPersistReplicas * kPersistReplicaStride;
int pcg_persist_capacity(int D, int device, bool batch = false);
hipError_t launch_pcg_persist(const Launch& L, int D, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
// the same for scenes of a batch: one cooperative launch of `blocks` workgroups over a block map of
// whole scenes (each scene's workgroups wait only on each other)
hipError_t launch_pcg_persist_batch(const SceneDev* tab, const MapEntry* map, int blocks, int D, hipStream_t stream,
                                    hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);

}  // namespace hfm
