// cdc_kernels.hip -- gfx950 kernels of the per-block Runge-Kutta update.
//
// One fused FP64 kernel per RK stage (stage_kernel_batch, over one run or a
// batch of runs) does, per tile of TX x TY cells inside one reference block:
//   phase 1  conserved -> primitive on the tile plus its 2-zone halo, halo
//            cells read straight from the neighbouring blocks' HBM with
//            periodic wrap (guard extension, src/scheme.rs:126-127), into LDS;
//            first failing cell recorded (try_block_primitive,
//            src/scheme.rs:185-199)
//   phase 2  PLM gradients gx, gy on the 1-zone ring (src/scheme.rs:224-225)
//   phase 3  x- and y-face fluxes: HLLE + viscous stress, constant-nu or
//            alpha viscosity, Plummer potential at the face
//            (src/scheme.rs:230-268, src/physics.rs:584-607, 776-803)
//   phase 4  flux divergence, source terms (gravity, sinks, buffer, fake mass,
//            cooling), the low-storage RK combine with the step base, and the
//            per-tile sums of the itemized source terms
//            (src/scheme.rs:273-316, src/state.rs:183-200)
// finalize_kernel_batch then sums each block's tiles in a fixed order, scales by
// dx*dy, applies the orbital-element perturbation and the RK combine of the
// block's integrated_source_terms / orbital_elements_change
// (src/scheme.rs:307-315, src/state.rs:195-197).
//
// FP64 reciprocals and square roots use the hardware estimates (v_rcp_f64,
// v_rsq_f64, about 5.6e-8 relative) refined by one Newton-Raphson step:
// measured over 524,288 log-uniform inputs in 1e-18 .. 1e18, 1/x within
// 10.8 ulp, 1/sqrt(x) within 19 ulp, sqrt(x) within 0.5 ulp of the exact
// values, at a fraction of the instructions of the IEEE expansions.  The
// max signal speed, which sets dt, uses IEEE division and sqrt.
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <limits.h>
#include <math.h>
#include <stdint.h>

#include "cdc_device.h"
#include "cdc_internal.h"

namespace cdc {

int nsum_for(int nv) { return nv == 4 ? SumMap<4>::n : SumMap<3>::n; }


// ---------------------------------------------------------------------------

// LDS doubles of one tile: primitives on the extended tile, the cell c_s^2
// plane (Euler), gradients on the 1-zone ring; the face fluxes and the tile
// reductions reuse the same space.
template <int NV, int TX, int TY, int NT>
struct TileLds {
    static constexpr int EXY = (TX + 4) * (TY + 4), RXY = (TX + 2) * (TY + 2);
    static constexpr int NR = SumMap<NV>::n + 1, CH = 16;
    static constexpr int A = NV * EXY + (NV == 4 ? EXY : 0) + 2 * NV * RXY;
    static constexpr int CHUNK_FIT = (A - NR * CH) / NT;
    static constexpr int CHUNK = CHUNK_FIT >= NR ? NR : (CHUNK_FIT < 1 ? 1 : CHUNK_FIT);
    static constexpr int R = NR * CH + CHUNK * NT;
    // coordinate table of the tile, kept apart from the reused space: x
    // vertices [XV, XV + TX], x centres, y vertices, y centres
    static constexpr int XV = 0, XC = TX + 1, YV = 2 * TX + 1, YC = 2 * TX + TY + 2, NGEO = 2 * (TX + TY) + 2;
    static constexpr int geo = A > R ? A : R;
    static constexpr int size = geo + NGEO;
};

// One RK stage on one TX x TY tile of one block (the P, F and U phases of
// try_advance_rk, src/scheme.rs:88-173, restricted to the tile), in the
// workgroup's LDS array lds[TileLds::size].
template <int NV, int TX, int TY, int NT>
__device__ __forceinline__ void stage_tile(const StageArgs &A, const int tile, double *lds, int &s_key) {
    constexpr int EX = TX + 4, EY = TY + 4, EXY = EX * EY;
    constexpr int RX = TX + 2, RY = TY + 2, RXY = RX * RY;
    constexpr int NFX = (TX + 1) * TY, NFY = TX * (TY + 1), NF = NFX + NFY;
    constexpr int KF = (NF + NT - 1) / NT;
    constexpr int CPT = TX * TY / NT;
    constexpr int NS = SumMap<NV>::n, NR = NS + 1;
    constexpr int CH = 16, CL = NT / CH;
    constexpr int NC = NV == 4 ? EXY : 0;       // cell c_s^2 plane (Euler alpha viscosity)
    constexpr int CHUNK = TileLds<NV, TX, TY, NT>::CHUNK;
    constexpr int NPASS = (NR + CHUNK - 1) / CHUNK;
    static_assert(NV * NF <= 2 * NV * RXY, "flux staging must fit the gradient region");
    static_assert((TX * TY) % NT == 0 && NT % CH == 0, "tile / thread mismatch");

    const int tid = threadIdx.x;
    const int bs = A.bs, N = A.N;
    const int tpby = bs / TY, tpb = (bs / TX) * tpby;
    const int b = tile / tpb, lt = tile - b * tpb;
    const int bi = b / A.nb, bj = b - (b / A.nb) * A.nb;
    const int li0 = (lt / tpby) * TX, lj0 = (lt - (lt / tpby) * tpby) * TY;
    const int gi0 = bi * bs + li0, gj0 = bj * bs + lj0;
    const size_t NN = (size_t)N * N;
    dev::gcdouble *guc = dev::gptr(A.uc), *gu0 = dev::gptr(A.u0), *gub = dev::gptr(A.ub), *gaux = dev::gptr(A.aux);
    dev::gdouble *guo = dev::gptr(A.uo);
    const dev::Axis axx = dev::block_axis(A.R, A.L, bi, bs);
    const dev::Axis axy = dev::block_axis(A.R, A.L, bj, bs);
    const bool alpha = A.k.viscosity_model == CDC_VISC_ALPHA;

    using TL = TileLds<NV, TX, TY, NT>;
    static_assert(TL::NGEO <= NT, "one thread per coordinate table entry");
    double *P = lds;
    double *C = lds + NV * EXY;
    double *GX = C + NC;
    double *GY = GX + NV * RXY;
    double *G = lds + TL::geo;

    // ---- phase 1: primitives on the tile + 2-zone halo --------------------
    int key = INT_MAX, kcode = 0;
    double kval = 0.0;
    if (tid == 0) s_key = INT_MAX;
    // all of this thread's halo-tile loads first, then the update phase's
    // operands, so that their latencies overlap each other and phases 1-3;
    // issued at raised wave priority, so that a starting tile's requests go
    // out ahead of the arithmetic of the other workgroups on the CU
    __builtin_amdgcn_s_setprio(3);
    constexpr int KE = (EXY + NT - 1) / NT;
    double ue[KE][NV];
#pragma unroll
    for (int q = 0; q < KE; ++q) {
        const int e = tid + q * NT;
        if (e < EXY) {
            const int ei = e / EY, ej = e - (e / EY) * EY;
            int gi = gi0 - 2 + ei, gj = gj0 - 2 + ej;
            gi = gi < 0 ? gi + N : (gi >= N ? gi - N : gi);
            gj = gj < 0 ? gj + N : (gj >= N ? gj - N : gj);
            const size_t g = (size_t)gi * N + gj;
#pragma unroll
            for (int v = 0; v < NV; ++v) ue[q][v] = guc[sidx(NV, v, NN, g)];
        }
    }
    double pu0[CPT][NV], pub[CPT][NV], pbr[CPT], pom[CPT];
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int c = tid + q * NT;
        const int ci = c / TY, cj = c - (c / TY) * TY;
        const size_t g = (size_t)(gi0 + ci) * N + (gj0 + cj);
        // read once per stage: streamed, so that L2 keeps the halo lines the
        // neighbouring tiles re-read
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            pu0[q][v] = __builtin_nontemporal_load(&gu0[sidx(NV, v, NN, g)]);
            pub[q][v] = A.read_base ? __builtin_nontemporal_load(&gub[sidx(NV, v, NN, g)]) : 0.0;
        }
        pbr[q] = __builtin_nontemporal_load(&gaux[g]);
        pom[q] = A.k.cooling_model == CDC_COOLING_BETA
                     ? __builtin_nontemporal_load(&gaux[(A.cool_factor ? 2 : 1) * NN + g])
                     : 0.0;
    }
    __builtin_amdgcn_s_setprio(0);
    // face and cell coordinates of the tile, once (bit-identical to the
    // per-face evaluation of the block's linspace)
    if (tid < TL::NGEO) {
        double x;
        if (tid < TL::XC) x = axx.v(li0 + tid);
        else if (tid < TL::YV) x = axx.c(li0 + tid - TL::XC);
        else if (tid < TL::YC) x = axy.v(lj0 + tid - TL::YV);
        else x = axy.c(lj0 + tid - TL::YC);
        G[tid] = x;
    }
#pragma unroll
    for (int q = 0; q < KE; ++q) {
        const int e = tid + q * NT;
        if (e < EXY) {
            const int ei = e / EY, ej = e - (e / EY) * EY;
            double p[NV], val = 0.0, id;
            const int code = dev::try_to_primitive<NV>(A.k, ue[q], p, val, id);
            if (code && ei >= 2 && ei < TX + 2 && ej >= 2 && ej < TY + 2) {
                const int c = (li0 + ei - 2) * bs + (lj0 + ej - 2);
                if (c < key) { key = c; kcode = code; kval = val; }
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) P[v * EXY + e] = p[v];
            if constexpr (NV == 4) {
                if (alpha) C[e] = A.k.gm * p[3] * id;
            }
        }
    }
    __syncthreads();

    // ---- phase 2: PLM gradients on the 1-zone ring --------------------------
    for (int r = tid; r < RXY; r += NT) {
        const int ri = r / RY, rj = r - (r / RY) * RY;
        const int ec = (ri + 1) * EY + (rj + 1);
        double thx = A.k.plm, thy = A.k.plm;
        if (A.k.has_pdt) {
            const double pdt = A.k.pdt;
            if (P[ec - EY] < pdt || P[ec] < pdt || P[ec + EY] < pdt) thx = 0.0;
            if (P[ec - 1] < pdt || P[ec] < pdt || P[ec + 1] < pdt) thy = 0.0;
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double *q = P + v * EXY + ec;
            GX[v * RXY + r] = dev::plm(thx, q[-EY], q[0], q[EY]);
            GY[v * RXY + r] = dev::plm(thy, q[-1], q[0], q[1]);
        }
    }
    __syncthreads();

    // ---- phase 3: face fluxes ----------------------------------------------
    // x-faces f < NFX, then y-faces; one code path for both kinds (the axis
    // enters by selects), so a wave holding both does not run both.
    double fl[KF][NV];
#pragma unroll
    for (int q = 0; q < KF; ++q) {
        const int f = tid + q * NT;
        if (f < NF) {
            const int isy = f >= NFX;
            const int g = isy ? f - NFX : f;
            const int fi = isy ? g / (TY + 1) : g / TY;
            const int fj = g - fi * (isy ? TY + 1 : TY);
            const int eL = isy ? (fi + 2) * EY + fj + 1 : (fi + 1) * EY + fj + 2;
            const int rL = isy ? (fi + 1) * RY + fj : fi * RY + fj + 1;
            const int de = isy ? 1 : EY, dr = isy ? 1 : RY;
            const double fxp = G[(isy ? TL::XC : TL::XV) + fi];
            const double fyp = G[(isy ? TL::YV : TL::YC) + fj];
            dev::face_flux<NV, EXY, RXY>(A, isy, P, C, GX, GY, eL, eL + de, rL, rL + dr, fxp, fyp, fl[q]);
        }
    }
    __syncthreads();
    double *FX = GX;
    double *FY = GX + NV * NFX;
#pragma unroll
    for (int q = 0; q < KF; ++q) {
        const int f = tid + q * NT;
        if (f < NFX) {
#pragma unroll
            for (int v = 0; v < NV; ++v) FX[v * NFX + f] = fl[q][v];
        } else if (f < NF) {
#pragma unroll
            for (int v = 0; v < NV; ++v) FY[v * NFY + f - NFX] = fl[q][v];
        }
    }
    __syncthreads();

    // ---- phase 4: update, sources, RK combine, tile sums ---------------------
    double acc[NR];
#pragma unroll
    for (int s = 0; s < NR; ++s) acc[s] = 0.0;
    const double ndt = -A.dt;
#pragma unroll
    for (int q = 0; q < CPT; ++q) {
        const int c = tid + q * NT;
        const int ci = c / TY, cj = c - (c / TY) * TY;
        const size_t g = (size_t)(gi0 + ci) * N + (gj0 + cj);
        double uc[NV], du[NV];
        const double *u0 = pu0[q], *ub = pub[q];
#pragma unroll
        for (int v = 0; v < NV; ++v) uc[v] = guc[sidx(NV, v, NN, g)];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const double *fx = FX + v * NFX + ci * TY + cj;
            const double *fy = FY + v * NFY + ci * (TY + 1) + cj;
            du[v] = ((fx[TY] - fx[0]) * A.k.idx + (fy[1] - fy[0]) * A.k.idy) * ndt;
        }
        // primitive recovery of u without floors (the sources' to_primitive)
        double vx, vy, id, p0 = 0.0;
        if (A.reuse_prim) {
            const int e = (ci + 2) * EY + (cj + 2);
            vx = P[1 * EXY + e];
            vy = P[2 * EXY + e];
            id = dev::frcp(uc[0]);
        } else {
            id = dev::frcp(uc[0]);
            vx = uc[1] * id;
            vy = uc[2] * id;
        }
        if constexpr (NV == 4) p0 = (uc[3] - 0.5 * uc[0] * (vx * vx + vy * vy)) * A.k.gm1;
        const double brate = pbr[q];
        const double cool = pom[q];
        double it[7][NV];
        dev::SinkGeom sg;
        dev::cell_sources<NV>(A, A.orb, uc, u0, G[TL::XC + ci], G[TL::YC + cj], vx, vy, id, p0, brate, cool, it, sg);
        double out[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            double total = it[CDC_ITEM_SINK1][v] + it[CDC_ITEM_SINK2][v] + it[CDC_ITEM_GRAV1][v] +
                           it[CDC_ITEM_GRAV2][v] + it[CDC_ITEM_BUFFER][v];
            if constexpr (NV == 4) total = total + it[CDC_ITEM_COOLING][v];
            else total = total + it[CDC_ITEM_FAKE_MASS][v];
            const double u1 = uc[v] + du[v] + total;
            out[v] = A.read_base ? u1 * A.c1 + ub[v] * A.c0 : u1;
            // streamed past L2: the next stage reads it back through its halo
            // loads, and dirty lines left in L2 cost more than they save
            __builtin_nontemporal_store(out[v], &guo[sidx(NV, v, NN, g)]);
        }
#pragma unroll
        for (int s = 0; s < NS; ++s) acc[s] += it[SumMap<NV>::item(s)][SumMap<NV>::comp(s)];
        if (A.want_amax) acc[NS] = fmax(acc[NS], dev::max_signal_speed<NV>(out, A.k.gm));
    }

    // ---- deterministic tile reductions --------------------------------------
    // Fixed-order two-level tree through LDS, in passes of CHUNK values so
    // that the reduction fits in the tile's own LDS.  Level 1: partial c of
    // value s sums entries c, c + CH, c + 2 CH, ... (consecutive lanes read
    // consecutive words: no bank conflicts).
    if (key != INT_MAX) atomicMin(&s_key, key);
    __syncthreads();
    double *red2 = lds;
    double *red = lds + NR * CH;
#pragma unroll
    for (int pass = 0; pass < NPASS; ++pass) {
        const int s0 = pass * CHUNK;
#pragma unroll
        for (int s = 0; s < NR; ++s)
            if (s >= s0 && s < s0 + CHUNK) red[(s - s0) * NT + tid] = acc[s];
        __syncthreads();
        const int nv_pass = (NR - s0) < CHUNK ? (NR - s0) : CHUNK;
        for (int w = tid; w < nv_pass * CH; w += NT) {
            const int sl = w / CH, c = w - (w / CH) * CH;
            const double *q = red + sl * NT + c;
            double t = q[0];
            if (s0 + sl == NS) {
                for (int j = 1; j < CL; ++j) t = fmax(t, q[j * CH]);
            } else {
                for (int j = 1; j < CL; ++j) t += q[j * CH];
            }
            red2[(s0 + sl) * CH + c] = t;
        }
        __syncthreads();
    }
    for (int s = tid; s < NR; s += NT) {
        const double *q = red2 + s * CH;
        double t = q[0];
        if (s == NS) {
            for (int j = 1; j < CH; ++j) t = fmax(t, q[j]);
        } else {
            for (int j = 1; j < CH; ++j) t += q[j];
        }
        dev::gptr(A.tile_sums)[(size_t)tile * NR + s] = t;
    }
    const int sk = s_key;
    if (sk == INT_MAX) {
        if (tid == 0) A.tile_err[tile].code = 0;
    } else if (key == sk) {
        TileErr te;
        te.code = kcode;
        te.cell = key;
        te.value = kval;
        A.tile_err[tile] = te;
        atomicCAS(A.err_flag, 0, A.stage_id);
    }
}

// Blocks of run r are global blocks begin[r] .. begin[r+1]-1; the offsets
// start at begin[0] = 0, so a single run needs no search and no load.
template <int NV>
__global__ void __launch_bounds__(64)
finalize_kernel_batch(const StageArgs *__restrict__ args, const int32_t *__restrict__ begin, int nruns) {
    __shared__ double scratch[kFinLds];
    const int r = nruns == 1 ? 0 : run_of(begin, nruns, blockIdx.x);
    const StageArgs &A = args[r];
    finalize_block<NV>(A, nruns == 1 ? (int)blockIdx.x : (int)blockIdx.x - begin[r], A.tpb, threadIdx.x, true,
                       scratch, scratch + 7 * 4, (int *)(scratch + 7 * 4 + 7 * 4));
}

// One RK stage of a batch of runs in one launch (a single run is a batch of
// one, so every run executes the same machine code): tiles of run r are
// global tiles begin[r] .. begin[r + 1] - 1, and each run reads its own
// StageArgs.  One-cell-per-thread tiles fit four 256-thread workgroups per
// CU in LDS; ask the register allocator for the matching four waves per SIMD.
// With fin, the first nfin workgroups finalize the earlier stage whose
// per-run arguments fin holds (blocks of run r: fin_begin[r] ..
// fin_begin[r + 1] - 1), one block per wave, ahead of the tiles; nfin is a
// multiple of 8, so the tiles keep their XCD mapping.  The two stages use
// different tile-sum buffers and the tiles read nothing the finalize writes.
template <int NV, int TX, int TY, int NT>
__global__ void __launch_bounds__(NT)
__attribute__((amdgpu_waves_per_eu((TX * TY == NT) ? 4 : 2)))
stage_kernel_batch(const StageArgs *__restrict__ args, const int32_t *__restrict__ begin, int nruns,
                   const StageArgs *__restrict__ fin, const int32_t *__restrict__ fin_begin, int fin_nruns,
                   int nfin) {
    static_assert((NT / 64) * kFinLds <= TileLds<NV, TX, TY, NT>::size, "finalize scratch fits the tile LDS");
    __shared__ double lds[TileLds<NV, TX, TY, NT>::size];
    __shared__ int s_key;
    if ((int)blockIdx.x < nfin) {
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int b = blockIdx.x * (NT / 64) + w;
        const bool active = b < fin_begin[fin_nruns];
        const int r = active ? run_of(fin_begin, fin_nruns, b) : 0;
        double *sc = lds + w * kFinLds;
        finalize_block<NV>(fin[r], active ? b - fin_begin[r] : 0, fin[r].tpb, lane, active, sc, sc + 7 * 4,
                           (int *)(sc + 7 * 4 + 7 * 4));
        return;
    }
    const int t = xcd_tile((int)blockIdx.x - nfin, begin[nruns]);
    const int r = run_of(begin, nruns, t);
    const StageArgs &A = args[r];
    if (stage_skipped(A)) return;
    stage_tile<NV, TX, TY, NT>(A, t - begin[r], lds, s_key);
}

// Batch prologue: clear every run's error flag (a kernel, not a memset, so
// that the sequence can be captured in a graph and replayed).
__global__ void __launch_bounds__(64) batch_reset_kernel(const StageArgs *__restrict__ args, int nruns) {
    for (int r = threadIdx.x; r < nruns; r += 64) *args[r].err_flag = 0;
}

// Batch epilogue, one workgroup per run: the run's max signal speed from the
// last stage's per-tile column, and its error flag.
__global__ void __launch_bounds__(256) batch_tail_kernel(const StageArgs *__restrict__ args, int nr,
                                                         double *amax, int32_t *flags) {
    __shared__ double red[256];
    const StageArgs &A = args[blockIdx.x];
    double m = 0.0;
    for (int i = threadIdx.x; i < A.ntiles; i += 256) m = fmax(m, A.tile_sums[(size_t)i * nr + nr - 1]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        amax[blockIdx.x] = red[0];
        flags[blockIdx.x] = *(volatile const int32_t *)A.err_flag;
    }
}

// Static per-cell planes (see AuxArgs), IEEE arithmetic in the reference's
// operation order.
__global__ void __launch_bounds__(256) aux_kernel(const AuxArgs a) {
#pragma clang fp contract(off)
    const int64_t N = a.N, NN = N * N;
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= NN) return;
    const int gi = (int)(c / N), gj = (int)(c - (c / N) * N);
    const dev::Axis ax = dev::block_axis(a.R, a.L, gi / a.bs, a.bs);
    const dev::Axis ay = dev::block_axis(a.R, a.L, gj / a.bs, a.bs);
    const double x = ax.c(gi % a.bs), y = ay.c(gj % a.bs);
    const double r = sqrt(x * x + y * y);
    const double yb = (r - a.R) / a.buffer_scale;
    a.aux[c] = 0.5 * a.buffer_rate * (1.0 + tanh(yb)) * a.omega_outer;
    if (a.naux > 1) {
        const double r2 = x * x + y * y + a.rs2;
        a.aux[NN + c] = sqrt(a.mtot / (r2 * sqrt(r2)));
    }
}

template <int NV>
__global__ void __launch_bounds__(256) max_speed_kernel(const double *u, int64_t ncell, double gm,
                                                        double *partial) {
    __shared__ double red[256];
    double m = 0.0;
    for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < ncell; c += (int64_t)gridDim.x * 256) {
        double w[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) w[v] = u[sidx(NV, v, ncell, c)];
        m = fmax(m, dev::max_signal_speed<NV>(w, gm));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = red[0];
}

// Boundary layout [block][i][j][v] <-> device layout (sidx)
__global__ void __launch_bounds__(256) aos_to_soa_kernel(int nv, int nb, int bs, const double *aos,
                                                         double *soa) {
    const int64_t N = (int64_t)nb * bs, NN = N * N;
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= NN) return;
    const int64_t bsq = (int64_t)bs * bs;
    const int64_t b = c / bsq, l = c - b * bsq;
    const int64_t i = l / bs, j = l - i * bs;
    const int64_t gi = (b / nb) * bs + i, gj = (b % nb) * bs + j;
    for (int v = 0; v < nv; ++v) soa[sidx(nv, v, NN, gi * N + gj)] = aos[c * nv + v];
}

__global__ void __launch_bounds__(256) soa_to_aos_kernel(int nv, int nb, int bs, const double *soa,
                                                         double *aos) {
    const int64_t N = (int64_t)nb * bs, NN = N * N;
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= NN) return;
    const int64_t bsq = (int64_t)bs * bs;
    const int64_t b = c / bsq, l = c - b * bsq;
    const int64_t i = l / bs, j = l - i * bs;
    const int64_t gi = (b / nb) * bs + i, gj = (b % nb) * bs + j;
    for (int v = 0; v < nv; ++v) aos[c * nv + v] = soa[sidx(nv, v, NN, gi * N + gj)];
}

// Re-gridding between two device states (sidx layout) of one domain, r a power of two:
// injection (up) or the mean of r x r fine cells as a pairwise tree over each
// row, then over the row means (down), the same order as the host resampler.
__global__ void __launch_bounds__(256) resample_kernel(int nv, const double *src, int64_t Ns, double *dst,
                                                       int64_t Nd, int r, int down) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= Nd * Nd) return;
    const int64_t gi = c / Nd, gj = c - (c / Nd) * Nd;
    const int64_t ns = Ns * Ns, nd = Nd * Nd;
    for (int v = 0; v < nv; ++v) {
        if (!down) {
            dst[sidx(nv, v, nd, c)] = src[sidx(nv, v, ns, (gi / r) * Ns + gj / r)];
            continue;
        }
        double rows[16];
        for (int a = 0; a < r; ++a) {
            double t[16];
            const int64_t q = (gi * r + a) * Ns + gj * r;
            for (int k = 0; k < r; ++k) t[k] = src[sidx(nv, v, ns, q + k)];
            for (int w = r; w > 1; w /= 2)
                for (int k = 0; k < w / 2; ++k) t[k] = t[2 * k] + t[2 * k + 1];
            rows[a] = t[0] / (double)r;
        }
        for (int w = r; w > 1; w /= 2)
            for (int k = 0; k < w / 2; ++k) rows[k] = rows[2 * k] + rows[2 * k + 1];
        dst[sidx(nv, v, nd, c)] = rows[0] / (double)r;
    }
}

__global__ void set_int_kernel(int32_t *p, int32_t v) { *p = v; }

// The beta-cooling relaxation factor exp(-dt Omega / beta) of every cell for
// an advance (one dt for all its stages), from aux plane 1 (Omega), the same
// expression the stage kernel evaluates when it is not given: into the march
// kernel's third pair plane (buffer rate, factor) when stat is given, stored
// whole (16 B per lane: a store of half of each pair leaves partly written
// lines), else into aux plane 2 (the tiled kernel's).
__global__ void __launch_bounds__(256) cooling_factor_kernel(const StageArgs *__restrict__ a, double *aux,
                                                              double *stat, int64_t nn) {
    const int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (c >= nn) return;
    const double dt = a->dt, cool = aux[nn + c];
    const double f = exp(-dt * cool * a->k.cooling_inv_beta);
    if (stat) reinterpret_cast<double2 *>(stat + 4 * nn)[c] = make_double2(aux[c], f);
    else aux[2 * nn + c] = f;
}

// the refined estimates, for measuring their error
__global__ void __launch_bounds__(256) math_probe_kernel(const double *x, double *out, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    out[3 * (size_t)i + 0] = dev::frcp(x[i]);
    out[3 * (size_t)i + 1] = dev::frsq(x[i]);
    out[3 * (size_t)i + 2] = dev::fsqrt(x[i]);
}

__global__ void reset_advance_kernel(int32_t *flag, double *amax) {
    *flag = 0;
    *amax = 0.0;
}

// max over a strided column (the per-tile max signal speed slot), one block
__global__ void __launch_bounds__(256) column_max_kernel(const double *x, int n, int stride, double *out) {
    __shared__ double red[256];
    double m = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) m = fmax(m, x[(size_t)i * stride]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// ---------------------------------------------------------------------------
// Host launchers

// Supported tile shapes (TX x TY cells, NT threads), largest first; the
// 64-column, 64-thread shapes are the column-marching kernel's strips.
static const TileShape kShapes[] = {{16, 32, 256}, {32, 16, 256}, {16, 16, 256}, {8, 32, 256},
                                    {32, 8, 256},  {8, 16, 128},  {16, 8, 128},  {8, 8, 64},
                                    {32, 64, 64},  {16, 64, 64}, {8, 64, 64}};

int tile_shapes(const TileShape **out) {
    *out = kShapes;
    return (int)(sizeof(kShapes) / sizeof(kShapes[0]));
}

TileShape tile_shape_for(int bs, int nv) {
    (void)nv;
    for (const TileShape &t : kShapes)
        if (bs % t.tx == 0 && bs % t.ty == 0) return t;
    return {8, 8, 64};
}

// With timing events, they are bound to the dispatch itself
// (hipExtLaunchKernelGGL): no marker packets between dependent launches.
template <int NV, int TX, int TY, int NT>
static hipError_t launch_batch_t(const StageArgs *args, const int32_t *begin, int nruns, int ntiles,
                                 const StageArgs *fin, const int32_t *fin_begin, int fin_nruns, int nblk,
                                 hipEvent_t ev0, hipEvent_t ev1, hipStream_t s) {
    constexpr int WPG = NT / 64;   // blocks finalized per workgroup
    const int nfin = fin ? ((nblk + WPG - 1) / WPG + 7) / 8 * 8 : 0;
    if (ev0 || ev1)
        hipExtLaunchKernelGGL((stage_kernel_batch<NV, TX, TY, NT>), dim3(ntiles + nfin), dim3(NT), 0, s, ev0, ev1, 0,
                              args, begin, nruns, fin, fin_begin, fin_nruns, nfin);
    else
        hipLaunchKernelGGL((stage_kernel_batch<NV, TX, TY, NT>), dim3(ntiles + nfin), dim3(NT), 0, s, args, begin,
                           nruns, fin, fin_begin, fin_nruns, nfin);
    return hipGetLastError();
}

template <int NV>
static hipError_t launch_batch_nv(const TileShape &ts, const StageArgs *args, const int32_t *begin, int nruns,
                                  int ntiles, const StageArgs *fin, const int32_t *fin_begin, int fin_nruns,
                                  int nblk, int cool, int visc, int read_base, hipEvent_t e0, hipEvent_t e1,
                                  hipStream_t s) {
    if (march_shape(ts))
        return launch_march_batch(NV, ts.tx, cool, visc, read_base, args, begin, nruns, ntiles, fin, fin_begin,
                                  fin_nruns, nblk, e0, e1, s);
#define CDC_SHAPE(X, Y, T)                                                                        \
    if (ts.tx == X && ts.ty == Y && ts.nt == T)                                                   \
        return launch_batch_t<NV, X, Y, T>(args, begin, nruns, ntiles, fin, fin_begin, fin_nruns, nblk, e0, e1, s);
    CDC_SHAPE(16, 32, 256)
    CDC_SHAPE(32, 16, 256)
    CDC_SHAPE(16, 16, 256)
    CDC_SHAPE(8, 32, 256)
    CDC_SHAPE(32, 8, 256)
    CDC_SHAPE(8, 16, 128)
    CDC_SHAPE(16, 8, 128)
    CDC_SHAPE(8, 8, 64)
#undef CDC_SHAPE
    return hipErrorInvalidValue;
}

hipError_t launch_stage_batch(int nv, const TileShape &ts, const StageArgs *args, const int32_t *tile_begin,
                              int nruns, int ntiles, const StageArgs *fin, const int32_t *block_begin, int nblk,
                              int cool, int visc, int read_base, hipEvent_t ev0, hipEvent_t ev1, hipStream_t s) {
    return nv == 3 ? launch_batch_nv<3>(ts, args, tile_begin, nruns, ntiles, fin, block_begin, nruns, nblk, cool,
                                        visc, read_base, ev0, ev1, s)
                   : launch_batch_nv<4>(ts, args, tile_begin, nruns, ntiles, fin, block_begin, nruns, nblk, cool,
                                        visc, read_base, ev0, ev1, s);
}

hipError_t launch_finalize_batch(int nv, const StageArgs *args, const int32_t *block_begin, int nruns,
                                 int nblocks, hipStream_t s) {
    if (nv == 3)
        hipLaunchKernelGGL(finalize_kernel_batch<3>, dim3(nblocks), dim3(64), 0, s, args, block_begin, nruns);
    else
        hipLaunchKernelGGL(finalize_kernel_batch<4>, dim3(nblocks), dim3(64), 0, s, args, block_begin, nruns);
    return hipGetLastError();
}

hipError_t launch_batch_reset(const StageArgs *args, int nruns, hipStream_t s) {
    hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(64), 0, s, args, nruns);
    return hipGetLastError();
}

hipError_t launch_batch_tail(int nv, const StageArgs *args, int nruns, double *amax, int32_t *flags,
                             hipStream_t s) {
    hipLaunchKernelGGL(batch_tail_kernel, dim3(nruns), dim3(256), 0, s, args, nsum_for(nv) + 1, amax, flags);
    return hipGetLastError();
}

hipError_t launch_aux(const AuxArgs &a, hipStream_t s) {
    const int64_t n = (int64_t)a.N * a.N;
    hipLaunchKernelGGL(aux_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

hipError_t launch_max_speed(int nv, const double *u, int64_t ncell, double gm, double *partial,
                            int nparts, hipStream_t s) {
    if (nv == 3) hipLaunchKernelGGL(max_speed_kernel<3>, dim3(nparts), dim3(256), 0, s, u, ncell, gm, partial);
    else hipLaunchKernelGGL(max_speed_kernel<4>, dim3(nparts), dim3(256), 0, s, u, ncell, gm, partial);
    return hipGetLastError();
}

hipError_t launch_aos_to_soa(int nv, int nb, int bs, const double *aos, double *soa, hipStream_t s) {
    const int64_t n = (int64_t)nb * bs * nb * bs;
    hipLaunchKernelGGL(aos_to_soa_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nv, nb, bs, aos, soa);
    return hipGetLastError();
}

hipError_t launch_soa_to_aos(int nv, int nb, int bs, const double *soa, double *aos, hipStream_t s) {
    const int64_t n = (int64_t)nb * bs * nb * bs;
    hipLaunchKernelGGL(soa_to_aos_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nv, nb, bs, soa, aos);
    return hipGetLastError();
}

hipError_t launch_column_max(const double *x, int n, int stride, double *out, hipStream_t s) {
    hipLaunchKernelGGL(column_max_kernel, dim3(1), dim3(256), 0, s, x, n, stride, out);
    return hipGetLastError();
}

hipError_t launch_resample(int nv, const double *src, int64_t Ns, double *dst, int64_t Nd, int r, int down,
                           hipStream_t s) {
    if (r < 1 || r > 16) return hipErrorInvalidValue;
    const int64_t n = Nd * Nd;
    hipLaunchKernelGGL(resample_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, nv, src, Ns, dst, Nd, r,
                       down);
    return hipGetLastError();
}

hipError_t launch_set_int(int32_t *p, int32_t v, hipStream_t s) {
    hipLaunchKernelGGL(set_int_kernel, dim3(1), dim3(1), 0, s, p, v);
    return hipGetLastError();
}

hipError_t launch_cooling_factor(const StageArgs *a, double *aux, double *stat, int64_t nn, hipStream_t s) {
    hipLaunchKernelGGL(cooling_factor_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, a, aux, stat,
                       nn);
    return hipGetLastError();
}

hipError_t launch_math_probe(const double *x, double *out, int n, hipStream_t s) {
    hipLaunchKernelGGL(math_probe_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, out, n);
    return hipGetLastError();
}

hipError_t launch_reset_advance(int32_t *flag, double *amax, hipStream_t s) {
    hipLaunchKernelGGL(reset_advance_kernel, dim3(1), dim3(1), 0, s, flag, amax);
    return hipGetLastError();
}

}  // namespace cdc
