// cdc_model.cpp -- the seeded cavity initial model (host).
//
// An InitialModel (src/traits.rs:105-116) with no reference counterpart: a
// disk with a cleared cavity and seeded noise.  It is written per cell over
// the block-decomposed mesh in the boundary layout [block][i][j][v], with
// cell centres from Mesh::cell_centers (src/mesh.rs:67-72) and conversion to
// conserved variables as Hydrodynamics::to_conserved.  The noise is a
// counter-based hash of (seed, stream, global cell index), so the state is a
// pure function of the global grid: identical for any block decomposition
// and regenerated bit for bit on restart (BlockData::from_model recomputes
// initial_conserved from the model, src/scheme.rs:46-57).
#include <math.h>
#include <stdint.h>

#include "../../include/cdc_solver.h"

namespace {

uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// uniform in [-1, 1)
double noise(uint64_t seed, uint64_t stream, uint64_t index) {
    const uint64_t key = splitmix64(seed ^ (stream * 0xD1B54A32D192ED03ull));
    const uint64_t h = splitmix64(key + index);
    return (double)(h >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

}  // namespace

extern "C" int32_t cdc_model_seeded_cavity(const cdc_mesh_t *mesh, const cdc_hydro_t *hydro,
                                           const cdc_cavity_model_t *m, double *out) {
    if (!mesh || !hydro || !m || !out) return CDC_ERR_INVALID_ARGUMENT;
    if (mesh->num_blocks < 1 || mesh->block_size < 1 || !(mesh->domain_radius > 0.0)) return CDC_ERR_INVALID_ARGUMENT;
    const int nb = mesh->num_blocks, bs = mesh->block_size;
    const int nv = hydro->kind == CDC_EULER ? 4 : 3;
    const int64_t N = (int64_t)nb * bs;
    const double R = mesh->domain_radius;
    const double L = 2.0 * R / (double)nb;
    const double gm = hydro->kind == CDC_EULER ? hydro->gamma_law_index : 1.0;
    const double rs2 = m->softening_length * m->softening_length;
    for (int bi = 0; bi < nb; ++bi) {
        for (int bj = 0; bj < nb; ++bj) {
            const double ax = -R + (double)bi * L, ay = -R + (double)bj * L;
            const double sx = ((ax + L) - ax) / (double)bs, sy = ((ay + L) - ay) / (double)bs;
            for (int i = 0; i < bs; ++i) {
                const double x = ((ax + sx * (double)i) + (ax + sx * (double)(i + 1))) / 2.0;
                for (int j = 0; j < bs; ++j) {
                    const double y = ((ay + sy * (double)j) + (ay + sy * (double)(j + 1))) / 2.0;
                    const uint64_t gidx = (uint64_t)((int64_t)(bi * bs + i) * N + (bj * bs + j));
                    const double r2 = x * x + y * y;
                    const double r = sqrt(r2);
                    double sd = r > 0.0 ? exp(-pow(m->cavity_radius / r, m->cavity_exponent)) : 0.0;
                    if (sd < m->density_floor) sd = m->density_floor;
                    sd = sd * (1.0 + m->noise_amplitude * noise(m->seed, 0, gidx));
                    const double t = r2 + rs2;
                    double vp = sqrt(r2 / (t * sqrt(t)));
                    if (m->taper_radius > 0.0) vp = vp / (1.0 + exp((r - m->taper_radius) / m->taper_width));
                    const double vx = r > 0.0 ? -vp * y / r : 0.0;
                    const double vy = r > 0.0 ? vp * x / r : 0.0;
                    double *u = out + (((int64_t)(bi * nb + bj) * bs + i) * bs + j) * nv;
                    u[0] = sd;
                    u[1] = sd * vx;
                    u[2] = sd * vy;
                    if (nv == 4) {
                        // cs^2 = -phi / Ma^2 of the softened unit mass at the centre
                        double p = sd * (1.0 / sqrt(t)) / (gm * m->mach_number * m->mach_number);
                        p = p * (1.0 + m->noise_amplitude * noise(m->seed, 1, gidx));
                        u[3] = 0.5 * sd * (vx * vx + vy * vy) + p / (gm - 1.0);
                    }
                }
            }
        }
    }
    return CDC_OK;
}
