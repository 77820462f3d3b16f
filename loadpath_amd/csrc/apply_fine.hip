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

// x-chunk of a fine-level kernel whose launch takes `units` blocks (or
// waves) per chunk: 16 node planes unless that leaves fewer than `target`
// of them, then 8 or 4 (a small grid would otherwise run a few dozen blocks,
// each marching its planes one after another)
int size_xchunk(int nx, int units, int target) {
    int xc = 16;
    while (xc > 4 && (long long)units * ((nx + xc - 1) / xc) < target) xc /= 2;
    return xc;
}

AGeom make_apply_geom(const Grid& Gd, int level) {
    const Level& lv = Gd.lv[level];
    AGeom G{};
    G.Nx = lv.g.Nx; G.Ny = lv.g.Ny; G.Nz = lv.g.Nz;
    G.nx = lv.g.nx; G.ny = lv.g.ny; G.nz = lv.g.nz;
    G.fnx = lv.fine.nx; G.fny = lv.fine.ny; G.fnz = lv.fine.nz;
    G.nn = lv.g.nn;
    G.ncell = lv.g.ncell;
    if (lv.kind == LK_FINE) {
        G.ny_per = 1;
        G.nyzb = (G.Ny * G.Nz + kBlock - 1) / kBlock;
        G.xchunk = Gd.tune_fine_xchunk > 0 ? Gd.tune_fine_xchunk : size_xchunk(G.Nx, G.nyzb, kSizeXChunkBlocks);
        G.nxc = (G.Nx + G.xchunk - 1) / G.xchunk;
        G.nblk = G.nyzb * G.nxc;
        G.per = (G.nblk + 7) / 8;
    } else {
        G.ny_per = 1;
        G.xchunk = 1;
        G.nblk = (int)((G.nn + (kBlock / 8) - 1) / (kBlock / 8));
        G.per = G.nblk;
    }
    return G;
}

int fine_vector_blocks(const Grid& Gd) { return 8 * make_apply_geom(Gd, 0).per; }

int apply_blocks(const Grid& Gd, int level, int eb) {
    if (half_explicit(Gd, level, eb)) return half_blocks(Gd, level);
    if (weights_mfma_path(Gd, level, eb)) return mfma_blocks(Gd, level);
    if (Gd.lv[level].kind == LK_FINE && eb == 4 && Gd.fp16()) return mfma_blocks(Gd, level);
    if (Gd.lv[level].kind == LK_EXPLICIT || explicit_weights(Gd, level, eb)) return explicit_blocks(Gd, level);
    const AGeom G = make_apply_geom(Gd, level);
    return Gd.lv[level].kind == LK_FINE ? 8 * G.per : G.nblk;
}

// algorithmic bytes of one operator launch: per right-hand side the node
// vectors the epilogue reads and writes (the operand included), once per
// launch the inverse diagonal the kinds that scale by it read (one vector
// shared by every right-hand side), the cell moduli (fine level) and the
// mask at one bit per dof (the per-apply unit of work's compulsory traffic)
static double apply_bytes(const Grid& G, int level, int eb, int nrhs, int epi) {
    int vecs = 2;          // per right-hand side
    bool dinv = false;     // plus one shared inverse diagonal
    switch (epi) {
        case EPI_CHEB: vecs = 6; dinv = true; break;          // d, x, r, dinv -> x, r, d2
        case EPI_CHEB_FIRST: vecs = 4; dinv = true; break;    // b, dinv -> x, r, d2
        case EPI_CHEB_ONLY: vecs = 2; dinv = true; break;     // b, dinv -> x
        case EPI_CHEB_LAST: vecs = 4; dinv = true; break;     // d, x, r, dinv -> x
        case EPI_CHEB_LAST_R: vecs = 3; dinv = true; break;   // r, x, dinv -> x
        case EPI_RESID_SCALED: vecs = 4; dinv = true; break;  // x, b, dinv -> r, d2
        case EPI_RESID_MASK: vecs = 3; break;                 // x, b -> out
        case EPI_JACOBI: vecs = 3; dinv = true; break;        // x, b, dinv -> out
        case EPI_POWER: vecs = 2; dinv = true; break;         // u, dinv -> out
        case EPI_RESID: vecs = 3; break;                      // x, b -> out
        default: vecs = 2; break;                             // u -> out
    }
    const Level& L = G.lv[level];
    const double v = 3.0 * (double)L.g.nn * eb;
    double b = (double)nrhs * vecs * v + (dinv ? v : 0.0) + 3.0 * (double)L.g.nn / 8.0;
    if (L.kind == LK_FINE) b += (double)L.g.ncell * eb;
    return b;
}

int launch_apply(const Grid& G, int level, int eb, const void* u, size_t vstride, int nrhs, const EpiArgs& E,
                 hipStream_t s) {
    Grid::Prof& P = G.prof;
    const bool rec = P.on && P.level == level && P.eb == eb;
    if (rec) {
        if (2 * (P.used + 1) > P.ev.size()) {
            P.dropped++;
        } else {
            LP_HIP(hipEventRecord(P.ev[2 * P.used], s));
        }
    }
    switch (G.lv[level].kind) {
        case LK_FINE:
            if (eb == 4 && G.fp16())
                launch_apply_fine_mfma(G, level, u, vstride, nrhs, E, s);
            else
                launch_apply_fine(G, level, eb, u, vstride, nrhs, E, s);
            break;
        case LK_WEIGHTS:
            if (half_explicit(G, level, eb))
                launch_apply_half(G, level, u, vstride, nrhs, E, s);
            else if (weights_mfma_path(G, level, eb))
                launch_apply_weights_mfma(G, level, u, vstride, nrhs, E, s);
            else if (explicit_weights(G, level, eb))
                launch_apply_explicit(G, level, eb, u, vstride, nrhs, E, s);
            else
                launch_apply_weights(G, level, eb, u, vstride, nrhs, E, s);
            break;
        default:
            if (half_explicit(G, level, eb))
                launch_apply_half(G, level, u, vstride, nrhs, E, s);
            else
                launch_apply_explicit(G, level, eb, u, vstride, nrhs, E, s);
            break;
    }
    if (rec && 2 * (P.used + 1) <= P.ev.size()) {
        LP_HIP(hipEventRecord(P.ev[2 * P.used + 1], s));
        P.bytes.push_back(apply_bytes(G, level, eb, nrhs, E.kind));
        P.used++;
    }
    return apply_blocks(G, level, eb);
}

}  // namespace lp
