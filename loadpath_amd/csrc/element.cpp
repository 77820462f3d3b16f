// Host-side FP64 element data: the unit-cube hex8 stiffness (fem.py:127-224)
// and the 8 child prolongations of a coarse cell (mg.py:57-73).
//
// Ke is integrated with 2x2x2 Gauss points (weight 1, det J = 1/8). The "im"
// element adds Wilson's nine incompatible modes (1 - x_m^2) per component and
// condenses them: Ke = Kuu - Kua Kaa^-1 Kua^T, then symmetrised, as
// fem._hex8_parts does (fem.py:191-217).
#include <cmath>
#include <cstring>

#include "lp_internal.h"

namespace lp {

namespace {

void strain_rows(const double grad[][3], int m, double* b /* 6 x 3m */) {
    const int cols = 3 * m;
    std::memset(b, 0, sizeof(double) * 6 * cols);
    for (int n = 0; n < m; ++n) {
        const double gx = grad[n][0], gy = grad[n][1], gz = grad[n][2];
        const int ux = 3 * n, uy = ux + 1, uz = ux + 2;
        b[0 * cols + ux] = gx;
        b[1 * cols + uy] = gy;
        b[2 * cols + uz] = gz;
        b[3 * cols + ux] = gy; b[3 * cols + uy] = gx;   // gamma_xy
        b[4 * cols + uy] = gz; b[4 * cols + uz] = gy;   // gamma_yz
        b[5 * cols + ux] = gz; b[5 * cols + uz] = gx;   // gamma_zx
    }
}

void isotropic(double nu, double c[6][6]) {
    const double lam = nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    const double mu = 1.0 / (2.0 * (1.0 + nu));
    std::memset(c, 0, sizeof(double) * 36);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) c[i][j] = lam;
    for (int i = 0; i < 3; ++i) {
        c[i][i] = lam + 2.0 * mu;
        c[3 + i][3 + i] = mu;
    }
}

// out (p x q) += w * A^T C B with A (6 x p), B (6 x q)
void add_atcb(double w, const double* a, int p, const double c[6][6], const double* b, int q, double* out) {
    double cb[6 * 24];
    for (int i = 0; i < 6; ++i)
        for (int j = 0; j < q; ++j) {
            double s = 0;
            for (int k = 0; k < 6; ++k) s += c[i][k] * b[k * q + j];
            cb[i * q + j] = s;
        }
    for (int i = 0; i < p; ++i)
        for (int j = 0; j < q; ++j) {
            double s = 0;
            for (int k = 0; k < 6; ++k) s += a[k * p + i] * cb[k * q + j];
            out[i * q + j] += w * s;
        }
}

}  // namespace

// trilinear (6 x 24) and bubble (6 x 9) strain operators of the unit cube at
// natural point nat (fem._b_matrix / _b_bubble, fem.py:127-189)
static void b_at(const double nat[3], double* b, double* ba) {
    double grad[8][3];
    for (int k = 0; k < 8; ++k) {
        const double s[3] = {(k & 1) ? 1.0 : -1.0, (k & 2) ? 1.0 : -1.0, (k & 4) ? 1.0 : -1.0};
        const double f0 = 1.0 + nat[0] * s[0], f1 = 1.0 + nat[1] * s[1], f2 = 1.0 + nat[2] * s[2];
        grad[k][0] = 0.25 * s[0] * f1 * f2;
        grad[k][1] = 0.25 * s[1] * f0 * f2;
        grad[k][2] = 0.25 * s[2] * f0 * f1;
    }
    double gb[3][3] = {{-4.0 * nat[0], 0, 0}, {0, -4.0 * nat[1], 0}, {0, 0, -4.0 * nat[2]}};
    strain_rows(grad, 8, b);
    strain_rows(gb, 3, ba);
}

// Ke (24 x 24) and, for the "im" element, the bubble recovery r (9 x 24):
// bubble amplitudes = r u (fem._hex8_parts, fem.py:191-217)
static void hex8_parts_host(double nu, int element, double* ke, double* r_out) {
    double c[6][6];
    isotropic(nu, c);
    const double g = 1.0 / std::sqrt(3.0);
    double kuu[24 * 24], kua[24 * 9], kaa[9 * 9];
    std::memset(kuu, 0, sizeof(kuu));
    std::memset(kua, 0, sizeof(kua));
    std::memset(kaa, 0, sizeof(kaa));
    for (int gp = 0; gp < 8; ++gp) {
        const double nat[3] = {(gp & 4) ? g : -g, (gp & 2) ? g : -g, (gp & 1) ? g : -g};
        double b[6 * 24], ba[6 * 9];
        b_at(nat, b, ba);
        add_atcb(0.125, b, 24, c, b, 24, kuu);
        add_atcb(0.125, b, 24, c, ba, 9, kua);
        add_atcb(0.125, ba, 9, c, ba, 9, kaa);
    }
    double k[24 * 24];
    std::memcpy(k, kuu, sizeof(k));
    if (element == 0) {
        // r = -Kaa^-1 Kua^T (9 x 24) by Gauss-Jordan on [Kaa | Kua^T]
        double m[9][9 + 24];
        for (int i = 0; i < 9; ++i) {
            for (int j = 0; j < 9; ++j) m[i][j] = kaa[i * 9 + j];
            for (int j = 0; j < 24; ++j) m[i][9 + j] = kua[j * 9 + i];
        }
        for (int col = 0; col < 9; ++col) {
            int piv = col;
            for (int i = col + 1; i < 9; ++i)
                if (std::fabs(m[i][col]) > std::fabs(m[piv][col])) piv = i;
            if (piv != col)
                for (int j = 0; j < 33; ++j) std::swap(m[col][j], m[piv][j]);
            const double d = m[col][col];
            for (int j = 0; j < 33; ++j) m[col][j] /= d;
            for (int i = 0; i < 9; ++i) {
                if (i == col) continue;
                const double f = m[i][col];
                if (f != 0.0)
                    for (int j = 0; j < 33; ++j) m[i][j] -= f * m[col][j];
            }
        }
        // k += Kua r, r = -m[:, 9:]
        for (int i = 0; i < 24; ++i)
            for (int j = 0; j < 24; ++j) {
                double s = 0;
                for (int a = 0; a < 9; ++a) s += kua[i * 9 + a] * (-m[a][9 + j]);
                k[i * 24 + j] += s;
            }
        if (r_out)
            for (int a = 0; a < 9; ++a)
                for (int j = 0; j < 24; ++j) r_out[a * 24 + j] = -m[a][9 + j];
    } else if (r_out) {
        std::memset(r_out, 0, sizeof(double) * 9 * 24);
    }
    for (int i = 0; i < 24; ++i)
        for (int j = 0; j < 24; ++j) ke[i * 24 + j] = 0.5 * (k[i * 24 + j] + k[j * 24 + i]);
}

void hex8_ke_host(double nu, int element, double* ke) { hex8_parts_host(nu, element, ke, nullptr); }

// C B at the stress sample points (constrained.py:383-387): the centre with
// the trilinear B alone (fem.hex8_b_center, fem.py:307-308), the 8 corners
// (x outermost, then y, then z), or the 8 quarter points (natural +-0.5,
// bit 0 = x; constrained.py:176-183), the last two with the bubbles' strain
// added (fem.hex8_b, fem.py:227-237). out (npts, 6, 24) at E = 1 and unit
// pitch.
void hex8_cb_host(double nu, int element, int points, double* out) {
    double c[6][6];
    isotropic(nu, c);
    double ke[24 * 24], r[9 * 24];
    hex8_parts_host(nu, element, ke, r);
    const int npts = points == 0 ? 1 : 8;
    for (int k = 0; k < npts; ++k) {
        double nat[3] = {0.0, 0.0, 0.0};
        if (points == 1) {
            nat[0] = (k & 4) ? 1.0 : -1.0;
            nat[1] = (k & 2) ? 1.0 : -1.0;
            nat[2] = (k & 1) ? 1.0 : -1.0;
        } else if (points == 2) {  // the quarter points, fem.CORNERS order (bit 0 = x)
            for (int d = 0; d < 3; ++d) nat[d] = ((k >> d) & 1) ? 0.5 : -0.5;
        }
        double b[6 * 24], ba[6 * 9];
        b_at(nat, b, ba);
        if (points != 0 && element == 0)
            for (int i = 0; i < 6; ++i)
                for (int j = 0; j < 24; ++j) {
                    double s = 0;
                    for (int a = 0; a < 9; ++a) s += ba[i * 9 + a] * r[a * 24 + j];
                    b[i * 24 + j] += s;
                }
        for (int i = 0; i < 6; ++i)
            for (int j = 0; j < 24; ++j) {
                double s = 0;
                for (int a = 0; a < 6; ++a) s += c[i][a] * b[a * 24 + j];
                out[(k * 6 + i) * 24 + j] = s;
            }
    }
}

void child_prolongations_host(double* pk) {
    std::memset(pk, 0, sizeof(double) * 8 * 576);
    for (int k = 0; k < 8; ++k)
        for (int q = 0; q < 8; ++q) {
            double pos[3];
            for (int d = 0; d < 3; ++d) pos[d] = 0.5 * (((k >> d) & 1) + ((q >> d) & 1));
            for (int r = 0; r < 8; ++r) {
                double w = 1.0;
                for (int d = 0; d < 3; ++d) w *= ((r >> d) & 1) ? pos[d] : 1.0 - pos[d];
                if (w != 0.0)
                    for (int i = 0; i < 3; ++i) pk[k * 576 + (3 * q + i) * 24 + 3 * r + i] = w;
            }
        }
}

}  // namespace lp

extern "C" int lp_hex8_cb(double nu, int element, int points, double* out) {
    if (!out || (element != 0 && element != 1) || (points < 0 || points > 2) || !(nu > -1.0 && nu < 0.5)) {
        lp::set_error("lp_hex8_cb: bad argument");
        return LP_ERR_ARG;
    }
    lp::hex8_cb_host(nu, element, points, out);
    return LP_OK;
}

extern "C" int lp_hex8_ke(double nu, int element, double* ke_out) {
    if (!ke_out || (element != 0 && element != 1) || !(nu > -1.0 && nu < 0.5)) {
        lp::set_error("lp_hex8_ke: bad argument");
        return LP_ERR_ARG;
    }
    lp::hex8_ke_host(nu, element, ke_out);
    return LP_OK;
}
