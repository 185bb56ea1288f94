// Host code under AddressSanitizer and UndefinedBehaviorSanitizer: the mesh
// readers and geometry builders of prandtl_amd/csrc/mesh.cpp and the oracle
// (oracle/prandtl_oracle.cpp), compiled from their sources with
// -fsanitize=address,undefined by tests/sanitize/Makefile (g++, no GPU code).
//
//   san_driver <mesh> ...   read each mesh file, refine a copy once, build the
//                           general-geometry tables at orders 1..3 (2-D) or
//                           2..3 (3-D); files the reader refuses must fail
//                           with PRANDTL_ERR_IO and a message
//   san_driver --oracle     oracle RHS and RK3 steps on small boxes: 2-D
//                           Euler with and without blending, 3-D
//                           Navier-Stokes faithful and corrected, a 1-D
//                           shock tube with slip walls
// Any sanitizer report aborts the process with a non-zero status.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/prandtl_hip.h"
#include "../../oracle/prandtl_oracle.h"
#include "../../prandtl_amd/csrc/mesh.h"

namespace {

// LGL nodes and weights on [0, 1] (Newton on (1-x^2) P_N'(x) on [-1, 1]) and
// the derivative matrix, for the geometry builders
void lgl(int N, std::vector<double> &x, std::vector<double> &w, std::vector<double> &D)
{
    const int Np = N + 1;
    std::vector<double> t(Np);
    for (int i = 0; i < Np; i++) {
        double z = -std::cos(M_PI * i / N);
        for (int it = 0; it < 100; it++) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= N; k++) {
                const double p2 = ((2 * k - 1) * z * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            const double dz = (z * p1 - p0) / (Np * p1);
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        t[i] = z;
    }
    x.resize(Np);
    w.resize(Np);
    for (int i = 0; i < Np; i++) {
        double p0 = 1.0, p1 = t[i];
        for (int k = 2; k <= N; k++) {
            const double p2 = ((2 * k - 1) * t[i] * p1 - (k - 1) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        x[i] = 0.5 * (t[i] + 1.0);
        w[i] = 1.0 / (N * Np * p1 * p1);
    }
    D.assign(Np * Np, 0.0);
    for (int i = 0; i < Np; i++)
        for (int j = 0; j < Np; j++) {
            if (i == j) continue;
            double li = 1.0, lj = 1.0;
            for (int k = 0; k < Np; k++) {
                if (k != i) li *= x[i] - x[k];
                if (k != j) lj *= x[j] - x[k];
            }
            D[i * Np + j] = (li / lj) / (x[i] - x[j]);
        }
    for (int i = 0; i < Np; i++) {
        double s = 0.0;
        for (int j = 0; j < Np; j++)
            if (j != i) s += D[i * Np + j];
        D[i * Np + i] = -s;
    }
}

int check_mesh(const char *path)
{
    prandtl_mesh_file *m = nullptr;
    const int rc = prandtl_mesh_read(path, &m);
    if (rc != PRANDTL_OK) {
        const char *msg = prandtl_mesh_last_error();
        std::printf("%s: refused (%d): %s\n", path, rc, msg ? msg : "");
        return (rc == PRANDTL_ERR_IO && msg && *msg) ? 0 : 1;
    }
    const int dim = prandtl_mesh_dim(m);
    for (int i = 0;; i++) {
        int attr = 0;
        const char *name = prandtl_mesh_name(m, i, &attr);
        if (!name) break;
    }
    int bad = 0;
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1 && prandtl_mesh_refine(m) != PRANDTL_OK) break;
        for (int N = (dim == 2 ? 1 : 2); N <= 3; N++) {
            std::vector<double> x, w, D;
            lgl(N, x, w, D);
            prandtl::quad::Geometry g;
            std::string err;
            bool ok;
            if (dim == 2) {
                prandtl_quad_mesh v = {};
                if (prandtl_mesh_view(m, &v) != PRANDTL_OK) { bad++; continue; }
                v.nattr = 0;
                for (int b = 0; b < v.nbdr; b++) v.nattr = v.bdr_attr[b] > v.nattr ? v.bdr_attr[b] : v.nattr;
                std::vector<int> bc(v.nattr, PRANDTL_BC_SLIP);
                std::vector<double> st(5 * (size_t)v.nattr, 1.0);
                v.attr_bc = bc.data();
                v.attr_state = st.data();
                ok = prandtl::quad::build(&v, N, x, w, D, false, g, err);
            } else {
                prandtl_hex_mesh v = {};
                if (prandtl_mesh_view_hex(m, &v) != PRANDTL_OK) { bad++; continue; }
                v.nattr = 0;
                for (int b = 0; b < v.nbdr; b++) v.nattr = v.bdr_attr[b] > v.nattr ? v.bdr_attr[b] : v.nattr;
                std::vector<int> bc(v.nattr, PRANDTL_BC_SLIP);
                std::vector<double> st(5 * (size_t)v.nattr, 1.0);
                v.attr_bc = bc.data();
                v.attr_state = st.data();
                ok = prandtl::quad::build_hex(&v, N, x, w, D, g, err);
            }
            std::printf("%s: dim %d refine %d N %d: %s (NE %d, hmin %.3g)\n", path, dim, pass, N,
                        ok ? "built" : err.c_str(), g.NE, g.hmin);
        }
    }
    prandtl_mesh_free(m);
    return bad;
}

int run_oracle()
{
    int bad = 0;
    auto box = [&](int dim, int N, int n, int parabolic, int blending, int quirks, const char *tag) {
        orc_params p = {};
        p.dim = dim;
        p.order = N;
        for (int r = 0; r < 3; r++) { p.nelem[r] = r < dim ? n : 1; p.lo[r] = 0.0; p.hi[r] = 1.0; }
        p.gamma = 1.4; p.Pr = 0.71; p.mu = parabolic ? 1e-3 : 0.0; p.parabolic = parabolic;
        p.blending = blending; p.alpha_max = 0.5; p.alpha_min = 0.001; p.quirks = quirks; p.threads = 1;
        if (dim == 1) { p.bc[0][0] = p.bc[0][1] = ORC_BC_SLIP; }
        orc_ctx *c = nullptr;
        if (orc_create(&p, &c) != 0) { std::printf("oracle %s: create failed\n", tag); bad++; return; }
        int NE, K, nv;
        orc_sizes(c, &NE, &K, &nv);
        std::vector<double> X((size_t)dim * NE * K), u((size_t)nv * NE * K), du(u.size());
        orc_coords(c, X.data());
        const size_t NK = (size_t)NE * K;
        for (size_t q = 0; q < NK; q++) {
            double s = 0.0;
            for (int d = 0; d < dim; d++) s += std::sin(2 * M_PI * X[d * NK + q] + d);
            const double rho = 1.0 + 0.2 * s + (blending && X[q] > 0.5 ? 0.5 : 0.0), p0 = 1.0 + 0.1 * s;
            double ke = 0.0;
            u[q] = rho;
            for (int d = 0; d < dim; d++) {
                const double v = dim == 1 ? 0.0 : 0.3 * std::cos(2 * M_PI * X[((d + 1) % dim) * NK + q]);
                u[(1 + d) * NK + q] = rho * v;
                ke += 0.5 * rho * v * v;
            }
            u[(dim + 1) * NK + q] = p0 / 0.4 + ke;
        }
        orc_rhs(c, u.data(), du.data());
        orc_rhs_ld(c, u.data(), du.data());
        for (int s = 0; s < 2; s++) orc_rk3_step(c, u.data(), 1e-4);
        std::vector<double> a(NE), e(NE);
        orc_alpha(c, a.data());
        orc_eta(c, e.data());
        double fl[8];
        orc_count_rhs_flops(c, u.data(), fl);
        double m = 0.0;
        for (double v : u) m = std::fmax(m, std::fabs(v));
        std::printf("oracle %s: NE %d K %d, max|u| %.4g, ops %.0f\n", tag, NE, K, m, fl[0]);
        if (!std::isfinite(m)) bad++;
        orc_destroy(c);
    };
    box(2, 3, 4, 0, 0, 1, "2-D Euler N3");
    box(2, 4, 4, 0, 1, 1, "2-D Euler blended N4");
    box(3, 3, 2, 1, 0, 1, "3-D Navier-Stokes faithful N3");
    box(3, 2, 2, 1, 1, 0, "3-D Navier-Stokes corrected blended N2");
    box(1, 3, 16, 0, 1, 1, "1-D slip walls blended N3");
    return bad;
}

} // namespace

int main(int argc, char **argv)
{
    int bad = 0;
    for (int i = 1; i < argc; i++) {
        if (std::strcmp(argv[i], "--oracle") == 0) bad += run_oracle();
        else bad += check_mesh(argv[i]);
    }
    std::printf(bad ? "FAIL %d\n" : "OK\n", bad);
    return bad ? 1 : 0;
}
