// Level operator instantiations for the "fine" cell-matrix kind (see apply_impl.h).
#include "apply_impl.h"

namespace lp {

void launch_apply_fine(const Grid& G, int level, int eb, const void* u, size_t vstride, int nrhs, const EpiArgs& E,
                     hipStream_t s) {
    if (eb == 8)
        apply_detail::launch_kind<double, LK_FINE>(G, level, static_cast<const double*>(u), vstride, nrhs, E, s);
    else
        apply_detail::launch_kind<float, LK_FINE>(G, level, static_cast<const float*>(u), vstride, nrhs, E, s);
}

}  // namespace lp

namespace lp {

AGeom make_apply_geom(const Grid& Gd, int level) {
    const Level& lv = Gd.lv[level];
    AGeom G{};
    G.Nx = lv.g.Nx; G.Ny = lv.g.Ny; G.Nz = lv.g.Nz;
    G.nx = lv.g.nx; G.ny = lv.g.ny; G.nz = lv.g.nz;
    G.fnx = lv.fine.nx; G.fny = lv.fine.ny; G.fnz = lv.fine.nz;
    G.nn = lv.g.nn;
    G.ncell = lv.g.ncell;
    G.xchunk = lv.kind == LK_FINE ? 8 : 4;
    G.nyzb = (G.Ny * G.Nz + kBlock - 1) / kBlock;
    G.nxc = (G.Nx + G.xchunk - 1) / G.xchunk;
    G.nblk = G.nyzb * G.nxc;
    G.per = (G.nblk + 7) / 8;
    return G;
}

int apply_blocks(const Grid& Gd, int level) { return 8 * make_apply_geom(Gd, level).per; }

int launch_apply(const Grid& G, int level, int eb, const void* u, size_t vstride, int nrhs, const EpiArgs& E,
                 hipStream_t s) {
    switch (G.lv[level].kind) {
        case LK_FINE: launch_apply_fine(G, level, eb, u, vstride, nrhs, E, s); break;
        case LK_WEIGHTS: launch_apply_weights(G, level, eb, u, vstride, nrhs, E, s); break;
        default: launch_apply_explicit(G, level, eb, u, vstride, nrhs, E, s); break;
    }
    return apply_blocks(G, level);
}

}  // namespace lp
