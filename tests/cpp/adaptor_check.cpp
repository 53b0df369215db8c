// C++ host program driving the engine through include/graeng.hpp with the
// reference's call shapes (Fwd::compute_G / G_vec_mul / GT_vec_mul and
// GaussNewtonInversion::solve_cg).  Reads a little-endian binary problem
// written by tests/test_gpu_cpp_adaptor.py and writes the results back.
//
//   in : int64 Nm, N_obs, flag, maxit; double lambda, tol;
//        prisms[Nm*6], points[N_obs*3], x[Nm], y[Nd], mk[Nm], mref[Nm],
//        Wd[Nd], dobs[Nd], P[Nm]; 4 x CSR {int64 n, nnz, rp[n+1]; int32 ci[nnz]; double v[nnz]}
//   out: G x [Nd], G^T y [Nm], dx [Nm], err, (double) iterations
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "graeng.hpp"

template <class T> static void rd(FILE *f, T *p, size_t n)
{
    if (fread(p, sizeof(T), n, f) != n) { fprintf(stderr, "short read\n"); exit(2); }
}

static graeng::Csr read_csr(FILE *f)
{
    graeng::Csr c;
    int64_t n, nnz;
    rd(f, &n, 1);
    rd(f, &nnz, 1);
    c.n = n;
    c.rp.resize(n + 1);
    c.ci.resize(nnz);
    c.v.resize(nnz);
    rd(f, c.rp.data(), n + 1);
    rd(f, c.ci.data(), nnz);
    rd(f, c.v.data(), nnz);
    return c;
}

static graeng::Csr diag(const std::vector<double> &d)
{
    graeng::Csr c;
    c.n = (int64_t)d.size();
    for (int64_t i = 0; i <= c.n; ++i) c.rp.push_back(i);
    for (int64_t i = 0; i < c.n; ++i) { c.ci.push_back((int32_t)i); c.v.push_back(d[i]); }
    return c;
}

int main(int argc, char **argv)
{
    if (argc != 3) { fprintf(stderr, "usage: adaptor_check in.bin out.bin\n"); return 2; }
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    int64_t hdr[4];
    double lt[2];
    rd(f, hdr, 4);
    rd(f, lt, 2);
    const int64_t Nm = hdr[0], N_obs = hdr[1];
    const unsigned long long flag = (unsigned long long)hdr[2];
    const int maxit = (int)hdr[3];
    int nc = 0;
    for (int b = 1; b < 10; ++b) nc += (flag >> b) & 1u;
    const int64_t Nd = nc * N_obs;
    std::vector<double> prisms(Nm * 6), points(N_obs * 3), x(Nm), y(Nd), mk(Nm), mref(Nm),
        Wd(Nd), dobs(Nd), P(Nm);
    rd(f, prisms.data(), prisms.size());
    rd(f, points.data(), points.size());
    rd(f, x.data(), Nm);
    rd(f, y.data(), Nd);
    rd(f, mk.data(), Nm);
    rd(f, mref.data(), Nm);
    rd(f, Wd.data(), Nd);
    rd(f, dobs.data(), Nd);
    rd(f, P.data(), Nm);
    graeng::Csr Ws = read_csr(f), Wx = read_csr(f), Wy = read_csr(f), Wz = read_csr(f);
    fclose(f);
    try {
        graeng::Context ctx(0);
        graeng::GaussNewtonInversion inv(ctx, prisms, points, flag);
        inv.compute_G();
        std::vector<double> Gx, GTy, norms;
        inv.G_vec_mul(x, Gx);
        inv.GT_vec_mul(y, GTy);
        inv.col_norms(norms);
        graeng::Csr I = diag(std::vector<double>(Nm, 1.0)), Wdm = diag(Wd), Pm = diag(P);
        double err = 0.0;
        int its = 0;
        std::vector<double> dx = inv.solve_cg(lt[1], maxit, err, its, lt[0], mk, mref, Wdm, dobs,
                                              Pm, Ws, Wx, Wy, Wz, I, I, I, I, I, I, I, I, I, I);
        FILE *o = fopen(argv[2], "wb");
        fwrite(Gx.data(), sizeof(double), Gx.size(), o);
        fwrite(GTy.data(), sizeof(double), GTy.size(), o);
        fwrite(dx.data(), sizeof(double), dx.size(), o);
        fwrite(norms.data(), sizeof(double), norms.size(), o);
        std::vector<double> Gx_c;
        inv.compute_G_corners();
        inv.G_vec_mul(x, Gx_c);
        fwrite(Gx_c.data(), sizeof(double), Gx_c.size(), o);
        double tail[2] = {err, (double)its};
        fwrite(tail, sizeof(double), 2, o);
        fclose(o);
    } catch (const graeng::Error &e) {
        fprintf(stderr, "graeng error: %s\n", e.what());
        return 1;
    }
    printf("adaptor_check ok: Nd=%lld Nm=%lld\n", (long long)Nd, (long long)Nm);
    return 0;
}
