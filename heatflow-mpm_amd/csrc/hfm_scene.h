// hfm_scene.h — scene state and kernel launchers of libhfm.so (host side).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/hfm.h"

namespace hfm {

// Device-resident control block of one PCG solve (warp.optim.linear.cg semantics).  Written by
// workgroup 0 of a kernel and read by later kernels only.
struct CgCtl {
    float slot_rr[2]; // ||r||^2 after iteration k, in slot k & 1
    float slot_rz[2]; // r.z after iteration k, in slot k & 1
    float atol_sq;    // fp32 threshold on ||r||^2
    int it;           // iterations done when the solve stopped
    int done;         // 1: converged or iteration cap reached; every later CG kernel returns at once
    int converged;
    int maxeff;       // iteration cap rounded up to the 2-iteration cycle
    int pad0;
    double tol;       // relative tolerance
    double atol;      // max(tol * ||b||, tol)
    float mean;       // pressure rhs mean to remove (numpy float32 mean)
    unsigned ticket2; // arrival counter of the rhs-mean chunks
};

// Pinned host memory written by the device with system-scope stores: CG progress and results,
// polled by the host to decide how far ahead to enqueue iterations.
struct CgMirror {
    int it;
    int done;
    int converged;
    int pad;
    double rr;
    double atol;
    double mean;
};

// Fixed launch geometry of a scene.
struct Geo {
    int n;
    long N;       // n^3 cells
    long F;       // (n+1) n^2 faces per axis
    float h;
};

struct Scene {
    hfm_params prm;
    Geo g;
    int device;
    hipStream_t stream;

    // air (face-major, each axis on its own lattice)
    float* u[3];          // velocity
    float* us[3];         // advected + buoyancy + relaxed velocity, projected in place, then swapped
    float* q[3];          // particle flux, zero on wall faces (solver.py:95-97, grid.py:62-67)
    float* kf[3];         // particle + static relaxation weights
    float* mf[3];         // particle + static relaxation momenta
    float* st_k[3];
    float* st_m[3];
    float* p;             // pressure (warm start across substeps)
    float* rhs;           // pressure rhs (mean removed)
    float* pdiag;         // pressure Jacobi preconditioner
    float* alpha_s;       // static + particle solid fraction
    float* st_alpha;
    int has_static;

    // heat
    float* T;
    float* T2;
    float* Tw;
    float* kap;
    float* alpha_p;       // particle-only solid fraction (heat.py:12-18)
    float* tsum;
    float* tfix;
    float* hrhs;
    float* hdiag;
    uint8_t* flags;       // FLAG_FIXED | FLAG_SOURCE per cell
    double rate, t_cap;

    // CG work vectors (the heat and pressure solves run one after the other)
    float* cg_r;
    float* cg_z;
    float* cg_d;          // search direction, two buffers of N (double-buffered by iteration parity)
    float* cg_Ap;
    double* partials;     // per-workgroup partial sums: [max_blocks] pass A, [3][max_blocks] pass B
    long max_blocks;
    CgCtl* ctl;           // [0] heat, [1] pressure (device)
    CgMirror* mirror;     // [0] heat, [1] pressure (pinned host)
    CgMirror* mirror_dev; // device view of mirror
    float* chunk_sums;    // numpy-mean chunk sums
    int* mean_prog;       // pairwise-sum programs: full chunk, then the tail chunk
    int mean_prog_tail;   // offset of the tail program in mean_prog, -1 if none
    float* scratch_field; // face-sized temporary for getters and diagnostics
    unsigned* umax;       // diagnostics maxima

    // particles
    long P;
    long P_cap;
    float* pos;           // (P, 3)
    float* vel;           // (P, 3)
    float* temp;
    float* melt;
    float* yield;
    float* visc;
    int* pkey;            // bin (cell) of each particle
    int* pslot;           // arrival slot inside its bin
    int* sorted;          // particle ids grouped by bin, ascending inside a bin
    int* bin_cnt;         // N + 1
    int* bin_start;       // N + 1 (exclusive scan)
    int* scan_tmp;
    int bins_dirty;

    hfm_stats stats;
    int pcg_cpt;          // minimum tiles per PCG workgroup (HFM_PCG_CPT)
    int flat_dir;         // 1: pressure pass A without LDS tiling (HFM_PCG_FLAT_DIR)
    int dir_ichunk;       // planes per workgroup of the tiled pass A, 0 = automatic (HFM_DIR_ICHUNK)
    int heat_pred;        // last exact heat iteration count (chunk sizing)
    int pres_pred;

    // kernel timing (HFM_FLAG_TIME_KERNELS)
    int timing;
    hipEvent_t* ev_pool;
    int ev_cap, ev_used;
    int* ev_class;
    double kt_ms[HFM_KT_COUNT];
    long long kt_n[HFM_KT_COUNT];
};

// ---- kernel launchers (hfm_fields.hip) ----
void launch_heat_prep(Scene* s, float dt);
void launch_advect_T(Scene* s, float dt);
void launch_g2p(Scene* s);
void launch_air_faces(Scene* s, float dt);
void launch_divergence(Scene* s, const float* const* vel, float scale, float* rhs_out, int with_diag);
void launch_rhs_mean(Scene* s, const int* prog_full, const int* prog_tail);
void launch_project(Scene* s, float dt_over_rho);
void launch_cell_velocity(Scene* s, float* out);
void launch_face_fraction(Scene* s, int axis, float* out);
void launch_abs_max(Scene* s, const float* a, long count, unsigned* slot);
void launch_apply_fixed(Scene* s, float* T);

// ---- kernel launchers (hfm_p2g.hip) ----
void launch_bin_particles(Scene* s);
void launch_p2g(Scene* s);

// ---- kernel launchers (hfm_pcg.hip); op 0 = heat, 1 = pressure ----
int pcg_blocks(const Scene* s);
void launch_pcg_init(Scene* s, int op, float dt);
void launch_pcg_dir(Scene* s, int op, float dt, int k, int report, hipEvent_t ev0 = nullptr,
                    hipEvent_t ev1 = nullptr);
void launch_pcg_upd(Scene* s, int op, int k, hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
int pcg_cpt(const Scene* s);

}  // namespace hfm
