// cdc_resample.cpp -- conservative re-gridding of a state between two meshes
// of one domain (host arrays in the boundary layout).
//
// Up-sampling is the reference checkpoint up-sampler's piecewise-constant
// injection (upsampler/src/main.rs:27-72: block_size doubled, num_blocks
// kept, conserved[i][j] = old[i/2][j/2], integrated_source_terms and
// orbital_elements_change carried over), generalised to any power-of-two
// ratio and to a change of num_blocks.  Down-sampling is its conservative
// inverse: each coarse cell is the mean of its r x r fine cells, summed as a
// pairwise tree (rows, then columns) so that an up-sampled field averages
// back to itself exactly.  Per-block ist / oec are copied when num_blocks is
// unchanged and summed over the merged blocks when coarse blocks cover whole
// groups of fine blocks; any other re-blocking zeroes them.
#include <string.h>

#include <string>
#include <vector>

#include "../../include/cdc_solver.h"
#include "cdc_host.h"

using cdc::host::set_msg;

namespace {

bool pow2(int64_t x) { return x > 0 && (x & (x - 1)) == 0; }

// global SoA index of boundary-layout element (block b, i, j)
inline int64_t gidx(int nb, int bs, int64_t b, int64_t i, int64_t j) {
    const int64_t N = (int64_t)nb * bs;
    return ((b / nb) * bs + i) * N + (b % nb) * bs + j;
}

}  // namespace

namespace cdc {
namespace host {

// mean of r consecutive values by a pairwise tree (r a power of two)
double tree_mean(const double *x, int64_t stride, int64_t r) {
    double buf[64] = {};
    std::vector<double> big;
    double *t = buf;
    if (r > 64) {
        big.resize((size_t)r);
        t = big.data();
    }
    for (int64_t k = 0; k < r; ++k) t[k] = x[k * stride];
    for (int64_t w = r; w > 1; w /= 2)
        for (int64_t k = 0; k < w / 2; ++k) t[k] = t[2 * k] + t[2 * k + 1];
    return t[0] / (double)r;
}

}  // namespace host
}  // namespace cdc

extern "C" {

int32_t cdc_resample_state(const cdc_mesh_t *src, const cdc_mesh_t *dst, int32_t nv, const double *u_src,
                           const double *ist_src, const double *oec_src, double *u_dst, double *ist_dst,
                           double *oec_dst) {
    if (!src || !dst || !u_src || !u_dst) return set_msg(CDC_ERR_INVALID_ARGUMENT, "null pointer");
    if (nv != 3 && nv != 4) return set_msg(CDC_ERR_INVALID_ARGUMENT, "nv must be 3 or 4");
    if (src->num_blocks < 1 || src->block_size < 1 || dst->num_blocks < 1 || dst->block_size < 1)
        return set_msg(CDC_ERR_INVALID_ARGUMENT, "empty mesh");
    if (src->domain_radius != dst->domain_radius)
        return set_msg(CDC_ERR_INVALID_ARGUMENT, "meshes must cover one domain");
    const int64_t Ns = (int64_t)src->num_blocks * src->block_size;
    const int64_t Nd = (int64_t)dst->num_blocks * dst->block_size;
    const bool down = Ns >= Nd;
    const int64_t r = down ? Ns / Nd : Nd / Ns;
    if ((down ? Ns % Nd : Nd % Ns) != 0 || !pow2(r))
        return set_msg(CDC_ERR_INVALID_ARGUMENT, "grid sizes must differ by a power-of-two factor");

    // global SoA planes of the source
    std::vector<double> s((size_t)(nv * Ns * Ns));
    const int64_t nbs = (int64_t)src->num_blocks * src->num_blocks, bss = src->block_size;
    for (int64_t b = 0; b < nbs; ++b)
        for (int64_t i = 0; i < bss; ++i)
            for (int64_t j = 0; j < bss; ++j) {
                const int64_t g = gidx(src->num_blocks, src->block_size, b, i, j);
                const double *e = u_src + ((b * bss + i) * bss + j) * nv;
                for (int v = 0; v < nv; ++v) s[(size_t)(v * Ns * Ns + g)] = e[v];
            }
    // destination planes
    std::vector<double> d((size_t)(nv * Nd * Nd));
    std::vector<double> row((size_t)r);
    for (int v = 0; v < nv; ++v) {
        const double *sp = s.data() + v * Ns * Ns;
        double *dp = d.data() + v * Nd * Nd;
        for (int64_t gi = 0; gi < Nd; ++gi)
            for (int64_t gj = 0; gj < Nd; ++gj) {
                if (!down) {
                    dp[gi * Nd + gj] = sp[(gi / r) * Ns + gj / r];
                    continue;
                }
                // mean over rows of the row means
                for (int64_t a = 0; a < r; ++a)
                    row[(size_t)a] = cdc::host::tree_mean(sp + (gi * r + a) * Ns + gj * r, 1, r);
                dp[gi * Nd + gj] = cdc::host::tree_mean(row.data(), 1, r);
            }
    }
    const int64_t nbd = (int64_t)dst->num_blocks * dst->num_blocks, bsd = dst->block_size;
    for (int64_t b = 0; b < nbd; ++b)
        for (int64_t i = 0; i < bsd; ++i)
            for (int64_t j = 0; j < bsd; ++j) {
                const int64_t g = gidx(dst->num_blocks, dst->block_size, b, i, j);
                double *e = u_dst + ((b * bsd + i) * bsd + j) * nv;
                for (int v = 0; v < nv; ++v) e[v] = d[(size_t)(v * Nd * Nd + g)];
            }

    cdc::host::resample_block_sums(src, dst, nv, ist_src, oec_src, ist_dst, oec_dst);
    return CDC_OK;
}

}  // extern "C"

void cdc::host::resample_block_sums(const cdc_mesh_t *src, const cdc_mesh_t *dst, int nv, const double *ist_src,
                                    const double *oec_src, double *ist_dst, double *oec_dst) {
    const bool down = (int64_t)src->num_blocks * src->block_size >= (int64_t)dst->num_blocks * dst->block_size;
    const int64_t nbd = (int64_t)dst->num_blocks * dst->num_blocks;
    const int64_t ni = 7 * nv, ne = 7 * 4;
    const int nbS = src->num_blocks, nbD = dst->num_blocks;
    if (ist_dst) memset(ist_dst, 0, (size_t)(nbd * ni) * sizeof(double));
    if (oec_dst) memset(oec_dst, 0, (size_t)(nbd * ne) * sizeof(double));
    if (nbS == nbD) {
        if (ist_dst && ist_src) memcpy(ist_dst, ist_src, (size_t)(nbd * ni) * sizeof(double));
        if (oec_dst && oec_src) memcpy(oec_dst, oec_src, (size_t)(nbd * ne) * sizeof(double));
    } else if (down && nbS % nbD == 0) {
        // each coarse block is a (nbS/nbD)^2 group of whole fine blocks
        const int g = nbS / nbD;
        for (int64_t B = 0; B < nbd; ++B) {
            const int64_t Bi = B / nbD, Bj = B % nbD;
            for (int a = 0; a < g; ++a)
                for (int c = 0; c < g; ++c) {
                    const int64_t b = (Bi * g + a) * nbS + (Bj * g + c);
                    if (ist_dst && ist_src)
                        for (int64_t k = 0; k < ni; ++k) ist_dst[B * ni + k] += ist_src[b * ni + k];
                    if (oec_dst && oec_src)
                        for (int64_t k = 0; k < ne; ++k) oec_dst[B * ne + k] += oec_src[b * ne + k];
                }
        }
    }
}
