// cdc_march.h -- the column-marching form of the fused stage kernel (the
// kernel template; cdc_march.hip, cdc_march16.hip and cdc_march8.hip
// instantiate it for strips of 32, 16 and 8 rows).
//
// One RK stage (the P, F and U phases of try_advance_rk, src/scheme.rs:88-173)
// over a strip of 64 columns x H rows inside one reference block (block_size
// a multiple of 64), one wave per strip.  Lane l owns column cj0 + l and
// marches along the strip's rows (x; down the block, or up it for every
// second row segment of a plain run): the primitives of the last three rows,
// the left state and stress factors of the previous row's x-face, that face's
// flux and the previous row's y-flux difference stay in registers, so every
// cell is converted to primitives once, its PLM gradients and stress factors
// are evaluated once, and every x- and y-face flux is evaluated once (the
// tiled kernel re-evaluates halo cells and recomputes both cells' stress at
// every face).  Neighbouring columns come from the adjacent lanes through DPP
// wavefront shifts.  The strip's y-halo (two columns on each side) is
// handled once per strip, before the march: lane q evaluates the y-face at
// one side of the strip for one row from the 8 cells around it and leaves the
// face flux and the halo column's primitives in LDS.  The next row's
// conserved state and the update's single-use operands are loaded one row
// ahead, so their latency overlaps the current row's arithmetic; no barrier
// is taken inside the march.
//
// The arithmetic is the reference's (src/scheme.rs:201-316, src/physics.rs:
// 322-340, 584-607, 776-803) with regroupings that differ from it only by
// rounding: a face's viscous flux is 0.5 f (A_L + A_R), A the cell's stress
// component times its density and the cell-local part of nu (constant nu,
// or alpha times the cell's c_s^2 for gamma-law gas), f the face-local
// factor (alpha's 1/sqrt(m1/q1^3 + m2/q2^3), times the face's c_s^2 for
// isothermal gas), evaluated once per face, in place of the reference's
// (sigma_L + sigma_R) / 2; the HLLE flux as U_L A + U_R B plus the pressure
// terms (march_face); the stress factors with folded coefficients.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

#include <type_traits>

#include "cdc_device.h"
#include "cdc_internal.h"

namespace cdc {

namespace {

// DPP wavefront shifts of a double (two 32-bit moves); a lane whose source
// lane lies outside the wave keeps `old`.
constexpr int kFromNext = 0x130;   // wave_shl:1, lane l receives lane l + 1's value
constexpr int kFromPrev = 0x138;   // wave_shr:1, lane l receives lane l - 1's value

template <int CTRL>
__device__ __forceinline__ double dpp(double old, double src) {
    const long long o = __double_as_longlong(old), v = __double_as_longlong(src);
    const int lo = __builtin_amdgcn_update_dpp((int)o, (int)v, CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(v >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// The same shift where the out-of-wave lane's value is never used: that lane
// receives 0 and no old value needs to be kept in the destination register.
template <int CTRL>
__device__ __forceinline__ double dpp0(double src) {
    const long long v = __double_as_longlong(src);
    const int lo = __builtin_amdgcn_mov_dpp((int)v, CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_mov_dpp((int)(v >> 32), CTRL, 0xF, 0xF, true);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// The stress coefficients of a stage: the normal components are
// (4/3 + lambda) d_n v_n + (lambda - 2/3) d_t v_t, with the 1/dx, 1/dy of
// the per-cell gradients folded in.
struct StressCoef {
    double nx, ty, ny, tx;   // (4/3 + l)/dx, (l - 2/3)/dy, (4/3 + l)/dy, (l - 2/3)/dx
    double vg;               // alpha / sqrt(gamma) times gamma (gamma-law gas, alpha viscosity)
    __device__ __forceinline__ explicit StressCoef(const StageConsts &k)
        : nx((4.0 / 3.0 + k.lam) * k.idx), ty((k.lam - 2.0 / 3.0) * k.idy), ny((4.0 / 3.0 + k.lam) * k.idy),
          tx((k.lam - 2.0 / 3.0) * k.idx), vg(k.visc * k.gm) {}
};

// Stress factors of one cell: Ax (normal-normal component on an x-face), Ay
// (on a y-face) and B (off-diagonal), each times Sigma and the cell-local
// part of nu (src/physics.rs:322-340, src/scheme.rs:230-247), the
// divergence term regrouped into the two coefficients of each normal
// component (rounding-level differences).  Gamma-law gas with alpha
// viscosity: Sigma (alpha / sqrt(gamma)) c_s^2, c_s^2 = gamma p / Sigma, is
// formed as p (alpha / sqrt(gamma)) gamma.
template <int NV>
__device__ __forceinline__ void stress_factors(const StageConsts &k, const StressCoef &sc, bool alpha,
                                               const double *p, const double *gx, const double *gy, double &ax,
                                               double &ay, double &b) {
    double K = p[0] * k.visc;
    if constexpr (NV == 4) {
        if (alpha) K = p[3] * sc.vg;
    }
    ax = K * (gx[1] * sc.nx + gy[2] * sc.ty);
    ay = K * (gy[2] * sc.ny + gx[1] * sc.tx);
    b = K * (gx[2] * k.idx + gy[1] * k.idy);
}

// A double held in a vector register from here on: the constants every face
// reads, so that they neither occupy scalar registers across the march nor
// need a copy into a vector register at each use (a vector instruction reads
// at most one scalar operand).
__device__ __forceinline__ double in_vgpr(double x) {
    double y;
    __asm__("v_mov_b64 %0, %1" : "=v"(y) : "v"(x));
    return y;
}

// The per-stage constants of the face evaluation, in vector registers.
struct FaceConsts {
    double x1, y1, x2, y2, m1, m2;   // the bodies (Orbit)
    double rs2, eps_dx, gm, inv_gm1, inv_mach2, plm;
    __device__ __forceinline__ explicit FaceConsts(const StageArgs &A)
        : x1(in_vgpr(A.orb.x1)), y1(in_vgpr(A.orb.y1)), x2(in_vgpr(A.orb.x2)), y2(in_vgpr(A.orb.y2)),
          m1(in_vgpr(A.orb.m1)), m2(in_vgpr(A.orb.m2)), rs2(in_vgpr(A.k.rs2)), eps_dx(in_vgpr(A.k.eps_dx)),
          gm(in_vgpr(A.k.gm)), inv_gm1(in_vgpr(A.k.inv_gm1)), inv_mach2(in_vgpr(A.k.inv_mach2)),
          plm(in_vgpr(A.k.plm)) {}
};

// One face (AXIS 0: x-face, 1: y-face) from its reconstructed states pl, pr:
// HLLE (Toro ch. 10) plus the viscous flux 0.5 f (sa, sb); (vbx, vby) the
// mean of the two cell-centre velocities (gamma-law energy flux).
template <int NV, int AXIS>
__device__ __forceinline__ void march_face(const FaceConsts &c, bool alpha, const double *pl,
                                           const double *pr, double vbx, double vby, double sa, double sb,
                                           double fxp, double fyp, double *F) {
    const double d1x = fxp - c.x1, d1y = fyp - c.y1;
    const double d2x = fxp - c.x2, d2y = fyp - c.y2;
    const double r1s = d1x * d1x + d1y * d1y;
    const double r2s = d2x * d2x + d2y * d2y;
    double cs2f = 0.0;
    if constexpr (NV == 3) cs2f = (c.m1 * dev::frsq(r1s + c.rs2) + c.m2 * dev::frsq(r2s + c.rs2)) * c.inv_mach2;
    double f = 0.5;
    if (alpha) {
        // 1 / sqrt(m1/q1^3 + m2/q2^3), q = r + 1e-12 dx (r = 0 gives q =
        // 1e-150 + 1e-12 dx, which rounds to 1e-12 dx as the reference's does)
        const double q1 = dev::fsqrt_nn(r1s) + c.eps_dx;
        const double q2 = dev::fsqrt_nn(r2s) + c.eps_dx;
        const double c1 = q1 * q1 * q1, c2 = q2 * q2 * q2;
        // rsq(X / Y) = Y rsq(X Y), X = m1 q2^3 + m2 q1^3, Y = q1^3 q2^3:
        // one transcendental instead of two
        const double Y = c1 * c2;
        const double itf = Y * dev::frsq((c.m1 * c2 + c.m2 * c1) * Y);
        f = NV == 3 ? 0.5 * (cs2f * itf) : 0.5 * itf;
    }
    const double tn = f * sa;
    const double to = f * sb;
    constexpr int NM = 1 + AXIS;   // the normal momentum component
    const double vnl = pl[NM], vnr = pr[NM];
    double ul[NV], ur[NV], csl, csr, ppl, ppr;
    ul[0] = pl[0]; ul[1] = pl[0] * pl[1]; ul[2] = pl[0] * pl[2];
    ur[0] = pr[0]; ur[1] = pr[0] * pr[1]; ur[2] = pr[0] * pr[2];
    if constexpr (NV == 3) {
        csl = csr = dev::fsqrt(cs2f);
        ppl = pl[0] * cs2f;
        ppr = pr[0] * cs2f;
    } else {
        ul[3] = 0.5 * pl[0] * (pl[1] * pl[1] + pl[2] * pl[2]) + pl[3] * c.inv_gm1;
        ur[3] = 0.5 * pr[0] * (pr[1] * pr[1] + pr[2] * pr[2]) + pr[3] * c.inv_gm1;
        const double gpl = c.gm * pl[3], gpr = c.gm * pr[3];
        csl = gpl * dev::frsq(gpl * pl[0]);   // sqrt(gamma p / Sigma)
        csr = gpr * dev::frsq(gpr * pr[0]);
        ppl = pl[3];
        ppr = pr[3];
    }
    const double ap = fmax(fmax(vnl + csl, vnr + csr), 0.0);
    const double am = fmin(fmin(vnl - csl, vnr - csr), 0.0);
    const double inv = dev::frcp(ap - am);
    // F = (ap F_L - am F_R - ap am (U_L - U_R)) / (ap - am) with the
    // weights wl = ap / (ap - am), wr = -am / (ap - am) and
    // wd = -ap am / (ap - am) formed once.  With F_L = U_L vnl + p_L e_n
    // (+ p_L vnl in the energy) this is U_L A + U_R B + the pressure terms,
    // A = wl vnl + wd, B = wr vnr - wd: two operations per component
    // (a regrouping, rounding-level differences).
    const double wl = ap * inv, wr = -am * inv, wd = -(ap * am) * inv;
    const double cA = wl * vnl + wd, cB = wr * vnr - wd;
    const double pwl = ppl * wl, pwr = ppr * wr;
#pragma unroll
    for (int v = 0; v < NV; ++v) F[v] = ul[v] * cA + ur[v] * cB;
    F[NM] += pwl + pwr;
    if constexpr (NV == 4) F[3] += pwl * vnl + pwr * vnr;
    const double tau_x = AXIS ? to : tn, tau_y = AXIS ? tn : to;
    F[1] -= tau_x;
    F[2] -= tau_y;
    if constexpr (NV == 4) F[3] -= tau_x * vbx + tau_y * vby;
}

// Buffer descriptors for the state arrays: one per array (4 SGPRs), the
// cell's byte offset in a VGPR and the plane's in an SGPR, so that the
// arrays' planes need no 64-bit base pointers of their own.  Offsets are
// 32-bit: the kernel runs on grids whose arrays hold under 2^32 bytes.
typedef __amdgpu_buffer_rsrc_t Rsrc;
constexpr int kBufWord3 = 0x00020000;   // raw buffer: 32-bit data format, no swizzle
constexpr int kNt = 2;                  // nontemporal cache policy

__device__ __forceinline__ Rsrc rsrc(const double *p, unsigned bytes) {
    return __builtin_amdgcn_make_buffer_rsrc((void *)p, (short)0, (int)bytes, kBufWord3);
}
template <int AUX = 0>
__device__ __forceinline__ double bload(Rsrc r, unsigned voff, unsigned soff) {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, AUX));
}
typedef double d2_t __attribute__((ext_vector_type(2)));
// two doubles at byte offset voff + soff (16-byte aligned): one 128-bit load
template <int AUX = 0>
__device__ __forceinline__ d2_t bload2(Rsrc r, unsigned voff, unsigned soff = 0u) {
    return __builtin_bit_cast(d2_t, __builtin_amdgcn_raw_buffer_load_b128(r, (int)voff, (int)soff, AUX));
}
// The stage's output stores take the default cache policy: the next stage
// reads them back (on the 2048^2 gamma-law workload 1.2% faster per launch
// than non-temporal stores on one box and the same on another, the whole
// step 0.3-0.5% faster on both; the single-use loads keep the non-temporal
// hint, which makes no difference there).
__device__ __forceinline__ void bstore(double x, Rsrc r, unsigned voff, unsigned soff) {
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(__attribute__((ext_vector_type(2))) unsigned, x), r,
                                          (int)voff, (int)soff, 0);
}

__device__ __forceinline__ void bstore2(double x, double y, Rsrc r, unsigned voff, unsigned soff) {
    __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(__attribute__((ext_vector_type(4))) unsigned,
                                                              d2_t{x, y}),
                                           r, (int)voff, (int)soff, 0);
}

// cell at byte offset voff, planes pl bytes apart (the strip-edge arrays)
template <int NV>
__device__ __forceinline__ void load_cell(Rsrc r, unsigned voff, unsigned pl, double *out) {
#pragma unroll
    for (int v = 0; v < NV; ++v) out[v] = bload(r, voff, (unsigned)v * pl);
}

// Cell c of a state in the sidx layout, g = 8 c and PB = 8 N^2 bytes: the
// pair (v0, v1) at 16 c, then (v2, v3) at 2 PB + 16 c, or v2 at 2 PB + 8 c
template <int NV, int AUX = 0>
__device__ __forceinline__ void load_state(Rsrc r, unsigned g, unsigned PB, double *out) {
    const d2_t a = bload2<AUX>(r, 2u * g);
    out[0] = a[0];
    out[1] = a[1];
    if constexpr (NV == 4) {
        const d2_t b = bload2<AUX>(r, 2u * g, 2u * PB);
        out[2] = b[0];
        out[3] = b[1];
    } else {
        out[2] = bload<AUX>(r, g, 2u * PB);
    }
}
template <int NV>
__device__ __forceinline__ void store_state(const double *x, Rsrc r, unsigned g, unsigned PB) {
    bstore2(x[0], x[1], r, 2u * g, 0u);
    if constexpr (NV == 4) bstore2(x[2], x[3], r, 2u * g, 2u * PB);
    else bstore(x[2], r, g, 2u * PB);
}

// PLM theta of the three cells a, b, c (plm_density_trigger); PDT = false:
// the run has no trigger
template <bool PDT>
__device__ __forceinline__ double theta3(const StageConsts &k, double plm, double da, double db, double dc) {
    if constexpr (PDT) {
        if (k.has_pdt && (da < k.pdt || db < k.pdt || dc < k.pdt)) return 0.0;
    }
    return plm;
}

template <int NV, int H>
struct MarchLds {
    static constexpr int HP = 0;                          // [2][H + 2][NV] halo-column primitives
    static constexpr int FB = 2 * (H + 2) * NV;           // [2][H][NV] strip-edge y-face fluxes
    static constexpr int ACC = FB + 2 * H * NV;           // [NS][64] per-lane sums of the itemized sources
    static constexpr int XC = ACC + SumMap<NV>::n * 64;   // [H] x of the strip's cell centres
    static constexpr int XV = XC + H;                     // [H + 1] x of the strip's vertices
    static constexpr int EO = XV + H + 1;                 // [4][H][NV] the result's edge columns 0, 1, 62, 63
    static constexpr int END = EO + 4 * H * NV;
    static constexpr int size = END > kFinLds ? END : kFinLds;
};

// One strip of one RK stage (see the file comment), tile = strip index in
// the run: strips are numbered block by block, row segments of H rows, then
// 64-column strips within a segment.  PLAIN: the run has neither a density
// floor nor a PLM density trigger, so the per-cell tests for them are left
// out and the update reuses the primitive recovery of its row (1/Sigma, the
// velocities and the pressure before the floor are the source terms'
// recovery of the same state).  UP: the strip marches from its last row to
// its first (see march_kernel_batch).
template <int NV, int H, int COOL, int RB, int VISC, bool PLAIN, bool UP>
__device__ __forceinline__ void march_strip(const StageArgs &A, const int tile, double *lds, int &s_key) {
    constexpr int NS = SumMap<NV>::n, NR = NS + 1;
    constexpr int NSINK = 2 * NV;   // the sink items lead the strip sums
    static_assert(SumMap<NV>::item(NSINK - 1) == CDC_ITEM_SINK2 && SumMap<NV>::item(NSINK) != CDC_ITEM_SINK2,
                  "sink items first");
    using ML = MarchLds<NV, H>;
    const int lane = threadIdx.x;
    const int bs = A.bs, N = A.N;
    const int spb = bs / 64;                          // strips across a block
    const int tpb = spb * (bs / H);
    const int b = tile / tpb, lt = tile - b * tpb;
    const int bi = b / A.nb, bj = b - (b / A.nb) * A.nb;
    const int ri0 = (lt / spb) * H, cj0 = (lt - (lt / spb) * spb) * 64;
    const size_t NN = (size_t)N * N;
    const unsigned PB = __builtin_amdgcn_readfirstlane((unsigned)(NN * 8));   // plane stride in bytes
    const Rsrc ruc = rsrc(A.uc, NV * PB), ruo = rsrc(A.uo, NV * PB);
    const Rsrc rub = rsrc(RB != 0 ? A.ub : A.uc, NV * PB);
    const Rsrc raux = rsrc(A.aux, (NV == 4 && COOL != kCoolT4 ? 3u : 1u) * PB);
    // the static operands in 16-byte pairs (StageArgs::stat), one descriptor
    // per pair plane
    constexpr bool kStat3 = NV == 4 && COOL == kCoolBetaFactor;
    const Rsrc rs0 = rsrc(A.stat, 2 * PB), rs1 = rsrc(A.stat + 2 * NN, 2 * PB);
    const Rsrc rs2 = rsrc(kStat3 ? A.stat + 4 * NN : A.stat, 2 * PB);
    const StageConsts &k = A.k;
    const dev::Axis axx = dev::block_axis(A.R, A.L, bi, bs);
    const dev::Axis axy = dev::block_axis(A.R, A.L, bj, bs);
    // the viscosity law, fixed at compile time (constant nu, nu = 0 included:
    // its stress terms are then zero)
    static_assert(VISC == kViscConst || VISC == kViscAlpha, "a march kernel is compiled for one viscosity law");
    constexpr bool alpha = VISC == kViscAlpha;
    const bool beta = k.cooling_model == CDC_COOLING_BETA;
    const int gi0 = bi * bs, gj0 = bj * bs;
    // wrapped global row / column of a block-local index
    auto grow = [&](int i) { const int g = gi0 + i; return g < 0 ? g + N : (g >= N ? g - N : g); };
    const FaceConsts fc(A);
    const StressCoef sc(k);
    // the orbit of the sources, positions and masses from the face constants
    Orbit ob = A.orb;
    ob.x1 = fc.x1; ob.y1 = fc.y1; ob.m1 = fc.m1;
    ob.x2 = fc.x2; ob.y2 = fc.y2; ob.m2 = fc.m2;
    double *hP = lds + ML::HP;
    double *Fb = lds + ML::FB;
    double *Acc = lds + ML::ACC;
    double *Eo = lds + ML::EO;

    int key = INT_MAX, kcode = 0;
    double kval = 0.0;
    if (lane == 0) s_key = INT_MAX;

    // ---- the strip's edges: y-faces at cj0 - 1/2 (side 0) and cj0 + 64 - 1/2
    // (side 1) for rows ri0 .. ri0 + H - 1, and the halo columns' primitives
    // for rows ri0 - 1 .. ri0 + H --------------------------------------------
    // The cells come from the input's strip-edge columns (StageArgs::ec): at
    // the boundary k = kL (side 0) or kL + 1 (side 1) the four columns
    // cL - 1, cL, cR, cR + 1 of a row are 32 contiguous bytes there, against
    // two 128-byte lines of the state per row and plane.
    const int NK = N / 64;
    const int kL = (gj0 + cj0) / 64, kR = kL + 1 == NK ? 0 : kL + 1;
    const unsigned PE = __builtin_amdgcn_readfirstlane((unsigned)NK * (unsigned)N * 32u);   // edge plane stride
    const Rsrc rec = rsrc(A.ec, NV * PE);
    for (int q = lane; q < 2 * H; q += 64) {
        const int s = q / H, r = ri0 + (q - s * H);
        const unsigned kk = s == 0 ? kL : kR;
        const int cR = s == 0 ? cj0 : cj0 + 64;   // block-local column right of the face
        const unsigned rm = ((unsigned)grow(r - 1) * NK + kk) * 4u, r0 = ((unsigned)grow(r) * NK + kk) * 4u,
                       rp = ((unsigned)grow(r + 1) * NK + kk) * 4u;
        // cells: 0 (r, cL-1), 1 (r, cL), 2 (r, cR), 3 (r, cR+1), 4 (r-1, cL), 5 (r+1, cL), 6 (r-1, cR), 7 (r+1, cR)
        const unsigned g8[8] = {r0, r0 + 1u, r0 + 2u, r0 + 3u, rm + 1u, rp + 1u, rm + 2u, rp + 2u};
        double u[8][NV];
#pragma unroll
        for (int c = 0; c < 8; ++c) load_cell<NV>(rec, g8[c] * 8u, PE, u[c]);
        double p[8][NV], id[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            double val;
            double praw;
            (void)dev::try_to_primitive<NV, !PLAIN>(k, u[c], p[c], val, id[c], praw);
        }
        double gxL[NV], gyL[NV], gxR[NV], gyR[NV];
        const double txL = theta3<!PLAIN>(k, fc.plm, p[4][0], p[1][0], p[5][0]);
        const double tyL = theta3<!PLAIN>(k, fc.plm, p[0][0], p[1][0], p[2][0]);
        const double txR = theta3<!PLAIN>(k, fc.plm, p[6][0], p[2][0], p[7][0]);
        const double tyR = theta3<!PLAIN>(k, fc.plm, p[1][0], p[2][0], p[3][0]);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            gxL[v] = dev::plm(txL, p[4][v], p[1][v], p[5][v]);
            gyL[v] = dev::plm(tyL, p[0][v], p[1][v], p[2][v]);
            gxR[v] = dev::plm(txR, p[6][v], p[2][v], p[7][v]);
            gyR[v] = dev::plm(tyR, p[1][v], p[2][v], p[3][v]);
        }
        double axL, ayL, bL, axR, ayR, bR;
        stress_factors<NV>(k, sc, alpha, p[1], gxL, gyL, axL, ayL, bL);
        stress_factors<NV>(k, sc, alpha, p[2], gxR, gyR, axR, ayR, bR);
        double pl[NV], pr[NV], F[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            pl[v] = p[1][v] + gyL[v] * 0.5;
            pr[v] = p[2][v] - gyR[v] * 0.5;
        }
        march_face<NV, 1>(fc, alpha, pl, pr, 0.5 * (p[1][1] + p[2][1]), 0.5 * (p[1][2] + p[2][2]),
                          ayL + ayR, bL + bR, axx.c(r), axy.v(cR), F);
        const int hr = r - ri0;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            Fb[(s * H + hr) * NV + v] = F[v];
            hP[(s * (H + 2) + hr + 1) * NV + v] = s == 0 ? p[1][v] : p[2][v];
            if (hr == 0) hP[(s * (H + 2)) * NV + v] = s == 0 ? p[4][v] : p[6][v];
            if (hr == H - 1) hP[(s * (H + 2) + H + 1) * NV + v] = s == 0 ? p[5][v] : p[7][v];
        }
        (void)axL; (void)axR;
    }

    // ---- prologue: march rows 0 .. 3 ------------------------------------------
    // The march visits block-local rows mrow(0), mrow(1), ...: ri0 - 2 ..
    // ri0 + H + 1 down the block, or ri0 + H + 1 .. ri0 - 2 up it.  The x-face
    // states of a row toward the row visited after it are its + side going
    // down and its - side going up (sg), and its PLM gradient takes its
    // neighbours in index order.
    auto mrow = [&](int n) { return UP ? ri0 + H + 1 - n : ri0 - 2 + n; };
    constexpr double sg = UP ? -0.5 : 0.5;
    const int j = cj0 + lane;                 // block-local column
    const int gj = gj0 + j;                   // global column (inside the block)
    const double yc = axy.c(j), yv1 = axy.v(j + 1);
    double Um2[NV], Um1[NV], U0r[NV], Uz[NV];
    load_state<NV>(ruc, ((unsigned)grow(mrow(0)) * N + gj) * 8u, PB, Um2);
    load_state<NV>(ruc, ((unsigned)grow(mrow(1)) * N + gj) * 8u, PB, Um1);
    load_state<NV>(ruc, ((unsigned)grow(mrow(2)) * N + gj) * 8u, PB, U0r);
    load_state<NV>(ruc, ((unsigned)grow(mrow(3)) * N + gj) * 8u, PB, Uz);
#pragma unroll
    for (int s = 0; s < NS; ++s) Acc[s * 64 + lane] = 0.0;
    // the rows' x coordinates, read by every step instead of evaluated
    double *Xc = lds + ML::XC, *Xv = lds + ML::XV;
    if (lane < H) Xc[lane] = axx.c(ri0 + lane);
    if (lane <= H) Xv[lane] = axx.v(ri0 + lane);
    __syncthreads();   // the edge data and coordinates in LDS

    // 1/Sigma and the pressure before the floor of march rows a and b
    double Pa[NV], Pb[NV], ida, idb, pra, prb;
    double XLa[NV], axa = 0.0, ba = 0.0;
    {
        double Pm2[NV], idm2, prm2, val;
        (void)dev::try_to_primitive<NV, !PLAIN>(k, Um2, Pm2, val, idm2, prm2);
        (void)dev::try_to_primitive<NV, !PLAIN>(k, Um1, Pa, val, ida, pra);
        const int code = dev::try_to_primitive<NV, !PLAIN>(k, U0r, Pb, val, idb, prb);
        if (code) { key = mrow(2) * bs + j; kcode = code; kval = val; }
        // march row 1 (a halo row): gradients, its x-face state toward row
        // 2, stress factors
        double gx[NV], gy[NV], PL[NV], PR[NV];
        const int ha = mrow(1) - ri0 + 1;   // halo-column entry of the row
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            PL[v] = dpp<kFromPrev>(hP[ha * NV + v], Pa[v]);
            PR[v] = dpp<kFromNext>(hP[((H + 2) + ha) * NV + v], Pa[v]);
        }
        const double tx = theta3<!PLAIN>(k, fc.plm, Pm2[0], Pa[0], Pb[0]);
        const double ty = theta3<!PLAIN>(k, fc.plm, PL[0], Pa[0], PR[0]);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            gx[v] = UP ? dev::plm(tx, Pb[v], Pa[v], Pm2[v]) : dev::plm(tx, Pm2[v], Pa[v], Pb[v]);
            gy[v] = dev::plm(ty, PL[v], Pa[v], PR[v]);
            XLa[v] = Pa[v] + gx[v] * sg;
        }
        double ay_unused;
        stress_factors<NV>(k, sc, alpha, Pa, gx, gy, axa, ay_unused, ba);
    }

    // ---- the march ------------------------------------------------------------
    // Step t (z = mrow(t + 3); b = mrow(t + 2), a = mrow(t + 1) the rows
    // visited before it): primitives of row z, gradients and stress factors
    // of row b, the x-face between rows a and b, the y-faces of row b
    // (t < H) and the update of row a (t >= 1).  (Going down: z = ri0 + t + 1,
    // b = z - 1, a = z - 2.)
    // Step 0 and step H are peeled off, so that every step of the loop issues
    // the same memory operations in the order their results are needed (the
    // next row's state, the next update's operands, this row's stores): the
    // vector memory counter retires loads and stores in order, and a wait for
    // the next row's state then never waits for this row's stores.
    // Carried: Pa, Pb = P(a), P(b); Ua, Ub, Uz = U(a), U(b), U(z); XLa, axa,
    // ba = row a's x-face state toward row b and its stress factors; Ga =
    // -dt (dFy/dy + s Fx(face before a)/dx) of row a (the update adds -dt s
    // Fx(face a, b)/dx; s = 1 going down, -1 going up); S* = the update
    // operands of row a.
    double Ga[NV], Ua[NV], Ub[NV];
#pragma unroll
    for (int v = 0; v < NV; ++v) { Ga[v] = 0.0; Ua[v] = Um1[v]; Ub[v] = U0r[v]; }
    double amax = 0.0;
    // -dt/dx (times s), -dt/dy: going down Ga = -dt (dFy/dy - Fx(- 1/2)/dx)
    // and the update adds -dt Fx(+ 1/2)/dx; going up the faces swap roles
    const double idxdt = k.idx * (UP ? A.dt : -A.dt), idydt = k.idy * -A.dt;
    double S0[NV], Sb[NV], Sr, Sc = 0.0;
    // update operands, one row ahead of their update (two rows ahead, in a
    // second register set, measured 6% slower): every step issues the same
    // loads (the cooling plane read from a valid plane when the law is
    // decided at run time)
    constexpr bool kReadCool = NV == 4 && COOL != kCoolT4;
    const unsigned cool_plane = __builtin_amdgcn_readfirstlane(beta ? (A.cool_factor ? 2u : 1u) * PB : 0u);
    auto load_update_operands = [&](int row) {
        const unsigned g = ((unsigned)(gi0 + row) * N + gj) * 8u;
        // u0 and the buffer rate (and the cooling factor) in 128-bit loads
        const d2_t q0 = bload2<kNt>(rs0, 2u * g), q1 = bload2<kNt>(rs1, 2u * g);
        S0[0] = q0[0];
        S0[1] = q0[1];
        S0[2] = q1[0];
        if constexpr (NV == 4) S0[3] = q1[1];
        else Sr = q1[1];
        if constexpr (RB != 0) load_state<NV, kNt>(rub, g, PB, Sb);
        if constexpr (kStat3) {
            const d2_t q2 = bload2<kNt>(rs2, 2u * g);
            Sr = q2[0];
            Sc = q2[1];
        } else if constexpr (NV == 4) {
            Sr = bload<kNt>(raux, g, 0u);
            if constexpr (kReadCool) Sc = bload<kNt>(raux, g, cool_plane);
        }
    };
    load_update_operands(mrow(2));

    auto step = [&](auto UPD, auto YF, auto NEXT, const int t) {
        constexpr bool upd = decltype(UPD)::value, yface = decltype(YF)::value, next = decltype(NEXT)::value;
        const int z = mrow(t + 3);
        const int hb = mrow(t + 2) - ri0;   // row b in the strip (0 .. H - 1 while t < H)
        const int ha = mrow(t + 1) - ri0;   // row a in the strip
        // next row's conserved state, one step ahead
        double Un[NV];
        if constexpr (next) load_state<NV>(ruc, ((unsigned)grow(mrow(t + 4)) * N + gj) * 8u, PB, Un);
        // primitives of row z
        double Pz[NV], idz, prz, val;
        const int code = dev::try_to_primitive<NV, !PLAIN>(k, Uz, Pz, val, idz, prz);
        if (code && z >= ri0 && z < ri0 + H && z * bs + j < key) { key = z * bs + j; kcode = code; kval = val; }
        // gradients of row b (the halo columns' primitives from LDS)
        double gx[NV], gy[NV], PL[NV], PR[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            PL[v] = dpp<kFromPrev>(hP[(hb + 1) * NV + v], Pb[v]);
            PR[v] = dpp<kFromNext>(hP[((H + 2) + hb + 1) * NV + v], Pb[v]);
        }
        const double tx = theta3<!PLAIN>(k, fc.plm, Pa[0], Pb[0], Pz[0]);
        const double ty = theta3<!PLAIN>(k, fc.plm, PL[0], Pb[0], PR[0]);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            gx[v] = UP ? dev::plm(tx, Pz[v], Pb[v], Pa[v]) : dev::plm(tx, Pa[v], Pb[v], Pz[v]);
            gy[v] = dev::plm(ty, PL[v], Pb[v], PR[v]);
        }
        double axb, ayb, bb;
        stress_factors<NV>(k, sc, alpha, Pb, gx, gy, axb, ayb, bb);
        // x-face between rows a and b (block-local vertex ri0 + t going
        // down, ri0 + H - t going up)
        double Fxb[NV];
        {
            double xb[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) xb[v] = Pb[v] - gx[v] * sg;
            march_face<NV, 0>(fc, alpha, UP ? xb : XLa, UP ? XLa : xb, 0.5 * (Pa[1] + Pb[1]),
                              0.5 * (Pa[2] + Pb[2]), axa + axb, ba + bb, Xv[UP ? H - t : t], yc, Fxb);
        }
        // y-faces of row b: lane l evaluates the face between columns
        // l and l + 1 from its neighbour's right-hand state; the strip's
        // last face and the face before its first column come from LDS
        double dFyb[NV];
        if constexpr (yface) {
            double yl[NV], yr[NV], F[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                yl[v] = Pb[v] + gy[v] * 0.5;
                const double own = Pb[v] - gy[v] * 0.5;
                yr[v] = dpp0<kFromNext>(own);
            }
            // lane 63's face is replaced by the strip-edge face below, so
            // what that lane receives from outside the wave is never used
            const double ayn = dpp0<kFromNext>(ayb), bn = dpp0<kFromNext>(bb);
            double vbx = 0.0, vby = 0.0;
            if constexpr (NV == 4) {
                vbx = 0.5 * (Pb[1] + PR[1]);
                vby = 0.5 * (Pb[2] + PR[2]);
            }
            march_face<NV, 1>(fc, alpha, yl, yr, vbx, vby, ayb + ayn, bb + bn, Xc[hb], yv1, F);
            // the strip-edge faces of this row, read by every lane (one
            // address: a broadcast) so that no lane-dependent branch waits
            // on LDS
            double fe0[NV], fe1[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                fe0[v] = Fb[hb * NV + v];
                fe1[v] = Fb[(H + hb) * NV + v];
            }
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double f = lane == 63 ? fe1[v] : F[v];
                dFyb[v] = f - dpp<kFromPrev>(fe0[v], f);
            }
        } else {
#pragma unroll
            for (int v = 0; v < NV; ++v) dFyb[v] = 0.0;
        }
        // update of row a (block-local row i)
        if constexpr (upd) {
            const int i = ri0 + ha;
            const unsigned g = ((unsigned)(gi0 + i) * N + gj) * 8u;
            double du[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) du[v] = Fxb[v] * idxdt + Ga[v];
            // primitive recovery of u without floors (the sources' to_primitive)
            double id, vx, vy, p0 = 0.0;
            if constexpr (PLAIN) {
                id = ida;
                vx = Pa[1];
                vy = Pa[2];
                if constexpr (NV == 4) p0 = pra;
            } else {
                id = dev::frcp(Ua[0]);
                if (A.reuse_prim) {
                    vx = Pa[1];
                    vy = Pa[2];
                } else {
                    vx = Ua[1] * id;
                    vy = Ua[2] * id;
                }
                if constexpr (NV == 4) p0 = (Ua[3] - 0.5 * Ua[0] * (vx * vx + vy * vy)) * k.gm1;
            }
            double it[7][NV];
            dev::SinkGeom sgeo;
            dev::cell_sources<NV, COOL, true>(A, ob, Ua, S0, Xc[ha], yc, vx, vy, id, p0, Sr, Sc, it, sgeo);
            const bool sinks = sgeo.sr1 != 0.0 || sgeo.sr2 != 0.0;
            // the other items into the strip sums (no return value: no
            // registers held), and their sum
            constexpr int LAST = NV == 4 ? CDC_ITEM_COOLING : CDC_ITEM_FAKE_MASS;
#pragma unroll
            for (int s = NSINK; s < NS; ++s)
                atomicAdd(&Acc[s * 64 + lane], it[SumMap<NV>::item(s)][SumMap<NV>::comp(s)]);
            double total[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v)
                total[v] = it[CDC_ITEM_GRAV1][v] + it[CDC_ITEM_GRAV2][v] + it[CDC_ITEM_BUFFER][v] + it[LAST][v];
            // the sink items only where a sink rate is nonzero (elsewhere they
            // are +-0), their geometry evaluated again there so that nothing of
            // it stays live through the other sources
            if (sinks) {
                const dev::SinkGeom sh = dev::sink_geom(k, ob, Xc[ha], yc);
                dev::sink_items<NV>(k, ob, A.dt, Ua, vx, vy, id, p0, sh, it);
#pragma unroll
                for (int v = 0; v < NV; ++v) total[v] = (it[CDC_ITEM_SINK1][v] + it[CDC_ITEM_SINK2][v]) + total[v];
#pragma unroll
                for (int s = 0; s < NSINK; ++s)
                    atomicAdd(&Acc[s * 64 + lane], it[SumMap<NV>::item(s)][SumMap<NV>::comp(s)]);
            }
            double out[NV];
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                const double u1 = Ua[v] + du[v] + total[v];
                if constexpr (RB != 0) out[v] = u1 * A.c1 + Sb[v] * A.c0;
                else out[v] = u1;
            }
            // the result's edge columns, for the next stage's strip edges
            if (lane < 2 || lane >= 62) {
                const int slot = lane < 2 ? lane : lane - 60;
#pragma unroll
                for (int v = 0; v < NV; ++v) Eo[(slot * H + ha) * NV + v] = out[v];
            }
            // the next update's operands, then this row's stores
            if constexpr (next) load_update_operands(mrow(t + 2));
            store_state<NV>(out, ruo, g, PB);
            // into the lane's own LDS slots (no return value: no registers held)
            if (A.want_amax) amax = fmax(amax, dev::max_signal_speed<NV>(out, k.gm));
        }
        // rotate the rows
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            XLa[v] = Pb[v] + gx[v] * sg;
            Pa[v] = Pb[v];
            Pb[v] = Pz[v];
            Ua[v] = Ub[v];
            Ub[v] = Uz[v];
            if constexpr (next) Uz[v] = Un[v];
            Ga[v] = dFyb[v] * idydt - Fxb[v] * idxdt;
        }
        ida = idb;
        idb = idz;
        pra = prb;
        prb = prz;
        axa = axb;
        ba = bb;
        // (the next row's state is left in flight across the loop's back
        // edge: forcing it complete at the end of every step measured the
        // same on 2048^2 Euler and 1.5% slower on 1024^2 isothermal)
    };

    const std::true_type yes;
    const std::false_type no;
    step(no, yes, yes, 0);
#pragma unroll 1
    for (int t = 1; t < H; ++t) step(yes, yes, yes, t);
    step(yes, no, no, H);

    // ---- the strip's sums over its cells: lane s adds up column s in lane
    // order; the max signal speed by a fixed butterfly --------------------------
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) amax = fmax(amax, __shfl_xor(amax, m, 64));
    __syncthreads();
    // the result's edge columns: per plane and row, columns 0, 1 at boundary
    // kL (entries 2, 3) and 62, 63 at kR (entries 0, 1), 32 contiguous bytes
    // unless kR wraps to 0
    if (A.eo) {
        for (int e = lane; e < 4 * H * NV; e += 64) {
            const int j = e & 3, row = (e >> 2) % H, v = (e >> 2) / H;
            const size_t at = ((size_t)v * N + (size_t)(gi0 + ri0 + row)) * NK;
            const size_t o = j < 2 ? (at + kL) * 4 + 2 + j : (at + kR) * 4 + (j - 2);
            dev::gptr(A.eo)[o] = Eo[(j * H + row) * NV + v];
        }
    }
    if (lane < NS) {
        const double *q = Acc + lane * 64;
        double x = q[0];
        for (int c = 1; c < 64; ++c) x += q[c];
        dev::gptr(A.tile_sums)[(size_t)tile * NR + lane] = x;
    } else if (lane == NS) {
        dev::gptr(A.tile_sums)[(size_t)tile * NR + NS] = amax;
    }
    // the strip's first failing cell (block-local C order)
    if (key != INT_MAX) atomicMin(&s_key, key);
    __syncthreads();
    const int sk = s_key;
    if (sk == INT_MAX) {
        if (lane == 0) A.tile_err[tile].code = 0;
    } else if (key == sk) {
        TileErr te;
        te.code = kcode;
        te.cell = key;
        te.value = kval;
        A.tile_err[tile] = te;
        atomicCAS(A.err_flag, 0, A.stage_id);
    }
}

// One RK stage of a batch of runs (a single run is a batch of one), one wave
// per strip; with fin, nfin workgroups after the strips finalize the earlier
// stage, one block each (see stage_kernel_batch).  Two waves per SIMD: the march
// keeps three rows of state and the next row's operands in registers.
template <int NV, int H, int COOL, int RB, int VISC>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
march_kernel_batch(const StageArgs *__restrict__ args, const int32_t *__restrict__ begin, int nruns,
                   const StageArgs *__restrict__ fin, const int32_t *__restrict__ fin_begin, int fin_nruns,
                   int nfin) {
    __shared__ double lds[MarchLds<NV, H>::size];
    __shared__ int s_key;
    const int ntiles = begin[nruns];
    if ((int)blockIdx.x >= ntiles) {
        // after the strips: every strip is resident from the start (the
        // register budget allows two waves per SIMD), these take the slots
        // the first strips free
        const int b = blockIdx.x - ntiles;
        const bool active = b < fin_begin[fin_nruns];
        const int r = active ? run_of(fin_begin, fin_nruns, b) : 0;
        finalize_block<NV>(fin[r], active ? b - fin_begin[r] : 0, fin[r].tpb, threadIdx.x, active, lds, lds + 7 * 4,
                           (int *)(lds + 7 * 4 + 7 * 4));
        return;
    }
    const int t = xcd_tile((int)blockIdx.x, ntiles);
    const int r = run_of(begin, nruns, t);
    const StageArgs &A = args[r];
    if (stage_skipped(A)) return;
    // Runs with neither a density floor nor a PLM trigger take the plain
    // march, and the row segments of their blocks alternate direction: the
    // two strips on either side of a segment boundary then read the rows
    // around it (one's own rows, the other's halo) at about the same time,
    // both at their first or both at their last steps, so that the second
    // read is an L2 hit (the strips of neighbouring blocks run on one XCD).
    const int lt = t - begin[r];
    if (A.k.has_density_floor || A.k.has_pdt) {
        march_strip<NV, H, COOL, RB, VISC, false, false>(A, lt, lds, s_key);
    } else {
        const int spb = A.bs / 64, seg = (lt % (spb * (A.bs / H))) / spb;
        if (seg & 1) march_strip<NV, H, COOL, RB, VISC, true, true>(A, lt, lds, s_key);
        else march_strip<NV, H, COOL, RB, VISC, true, false>(A, lt, lds, s_key);
    }
}

template <int NV, int H, int COOL, int RB, int VISC>
hipError_t launch_march_t(const StageArgs *args, const int32_t *begin, int nruns, int ntiles, const StageArgs *fin,
                          const int32_t *fin_begin, int fin_nruns, int nblk, hipEvent_t ev0, hipEvent_t ev1,
                          hipStream_t s) {
    const int nfin = fin ? nblk : 0;   // one block per workgroup
    if (ev0 || ev1)
        hipExtLaunchKernelGGL((march_kernel_batch<NV, H, COOL, RB, VISC>), dim3(ntiles + nfin), dim3(64), 0, s, ev0, ev1, 0,
                              args,
                              begin, nruns, fin, fin_begin, fin_nruns, nfin);
    else
        hipLaunchKernelGGL((march_kernel_batch<NV, H, COOL, RB, VISC>), dim3(ntiles + nfin), dim3(64), 0, s, args, begin,
                           nruns, fin,
                           fin_begin, fin_nruns, nfin);
    return hipGetLastError();
}


// The launch for strips of H rows: the kernel compiled for the stage's
// cooling law, viscosity law and whether it reads the RK step base.
template <int H>
hipError_t launch_march_h(int nv, int cool, int visc, int read_base, const StageArgs *args, const int32_t *begin,
                          int nruns, int ntiles, const StageArgs *fin, const int32_t *fin_begin, int fin_nruns,
                          int nblk, hipEvent_t ev0, hipEvent_t ev1, hipStream_t s) {
    const int c = nv == 3 ? (int)kCoolAny : cool;
    const int rb = read_base ? 1 : 0;
#define CDC_MARCH_V(V, C, B, W)                                                                                 \
    if (nv == V && c == C && rb == B && visc == W)                                                            \
        return launch_march_t<V, H, C, B, W>(args, begin, nruns, ntiles, fin, fin_begin, fin_nruns, nblk, ev0, ev1, \
                                             s);
#define CDC_MARCH(V, C, B) CDC_MARCH_V(V, C, B, kViscConst) CDC_MARCH_V(V, C, B, kViscAlpha)
#define CDC_MARCH_RB(V, C) CDC_MARCH(V, C, 0) CDC_MARCH(V, C, 1)
    CDC_MARCH_RB(3, kCoolAny)
    CDC_MARCH_RB(4, kCoolT4)
    CDC_MARCH_RB(4, kCoolBetaFactor)
    CDC_MARCH_RB(4, kCoolBetaExp)
#undef CDC_MARCH_RB
#undef CDC_MARCH
#undef CDC_MARCH_V
    return hipErrorInvalidValue;
}

}  // namespace

// Strips of 16 rows (cdc_march16.hip) and of 8 rows (cdc_march8.hip).
hipError_t launch_march_rows16(int nv, int cool, int visc, int read_base, const StageArgs *args,
                               const int32_t *begin, int nruns, int ntiles, const StageArgs *fin,
                               const int32_t *fin_begin, int fin_nruns, int nblk, hipEvent_t ev0, hipEvent_t ev1,
                               hipStream_t s);
hipError_t launch_march_rows8(int nv, int cool, int visc, int read_base, const StageArgs *args,
                              const int32_t *begin, int nruns, int ntiles, const StageArgs *fin,
                              const int32_t *fin_begin, int fin_nruns, int nblk, hipEvent_t ev0, hipEvent_t ev1,
                              hipStream_t s);

}  // namespace cdc
