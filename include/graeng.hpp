// C++ adaptor over the graeng C ABI with the reference's member-function
// shapes, header-only and free of any GPU-runtime or Eigen dependency.
//
//   graeng::Fwd                 <-> class Fwd               (src/Forward_modelling/Fwd.h)
//     compute_G()                    Fwd::compute_G          (Fwd.cpp:47-79)
//     compute_G_rectilinear(xs,ys,zs)  the same for tensor-product meshes (corner sums)
//     compute_G_corners()            the same for any mesh (per-prism corner sums)
//     G_vec_mul(vec, product)        Fwd::G_vec_mul          (Fwd.cpp:114-147)
//     GT_vec_mul(vec, product)       Fwd::GT_vec_mul         (Fwd.cpp:80-113)
//     compute_gobs(rho)              Fwd::compute_gobs       (Fwd.h:23-27)
//     col_norms(norms)               G.col(i).norm() reads of
//                                    InversionBase::set_depth_weighting(beta, 1)
//                                    (InversionBase.cpp:519-526)
//   graeng::GaussNewtonInversion::solve_cg
//                               <-> GaussNewtonInversion::solve_cg (GaussNewtonInversion.h:70-93)
//
// Vector arguments may be any type with data()/size()/resize() over doubles
// (std::vector<double>, Eigen::VectorXd).  Sparse operators may be any
// compressed row-major matrix exposing rows(), nonZeros(), outerIndexPtr(),
// innerIndexPtr(), valuePtr() (Eigen::SparseMatrix<double, Eigen::RowMajor>,
// the reference's SMatrix, gs.h:29) or a graeng::Csr.  Diagonal operators
// (Wd, P, WL_*) are passed as such matrices too, as in the reference, and
// their diagonals are extracted.  Errors throw graeng::Error (the reference
// asserts/aborts instead).
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "graeng.h"

namespace graeng {

struct Error : std::runtime_error {
    gra_status status;
    Error(gra_status s, const std::string &m) : std::runtime_error(m), status(s) {}
};

inline void check(gra_ctx *c, gra_status s, const char *where)
{
    if (s != GRA_OK)
        throw Error(s, std::string(where) + ": " + gra_status_string(s) + " " +
                           (c ? gra_last_error(c) : ""));
}

// Owned CSR copy with int64 row pointers.
struct Csr {
    int64_t n = 0;
    std::vector<int64_t> rp;
    std::vector<int32_t> ci;
    std::vector<double> v;
    gra_csr view() const
    {
        return gra_csr{n, (int64_t)v.size(), rp.data(), ci.data(), v.data()};
    }
    // the compressed row-major matrix interface used by to_csr / diagonal
    int64_t rows() const { return n; }
    int64_t nonZeros() const { return (int64_t)v.size(); }
    const int64_t *outerIndexPtr() const { return rp.data(); }
    const int32_t *innerIndexPtr() const { return ci.data(); }
    const double *valuePtr() const { return v.data(); }
};

template <class SM> Csr to_csr(const SM &m)
{
    Csr c;
    c.n = (int64_t)m.rows();
    c.rp.resize((size_t)c.n + 1);
    const auto *op = m.outerIndexPtr();
    for (int64_t i = 0; i <= c.n; ++i) c.rp[(size_t)i] = (int64_t)op[i];
    int64_t nnz = c.rp[(size_t)c.n];
    c.ci.assign(m.innerIndexPtr(), m.innerIndexPtr() + nnz);
    c.v.assign(m.valuePtr(), m.valuePtr() + nnz);
    return c;
}
inline Csr to_csr(const Csr &c) { return c; }

// Diagonal of a row-major sparse matrix (identity rows without an entry: 0).
template <class SM> std::vector<double> diagonal(const SM &m)
{
    int64_t n = (int64_t)m.rows();
    std::vector<double> d((size_t)n, 0.0);
    const auto *op = m.outerIndexPtr();
    const auto *ip = m.innerIndexPtr();
    const double *vp = m.valuePtr();
    for (int64_t i = 0; i < n; ++i)
        for (auto k = op[i]; k < op[i + 1]; ++k)
            if ((int64_t)ip[k] == i) d[(size_t)i] += vp[k];
    return d;
}
inline std::vector<double> diagonal(const std::vector<double> &d) { return d; }

class Context {
public:
    explicit Context(int device = 0) { check(nullptr, gra_ctx_create(device, &c_), "gra_ctx_create"); }
    ~Context() { gra_ctx_destroy(c_); }
    Context(const Context &) = delete;
    Context &operator=(const Context &) = delete;
    gra_ctx *get() const { return c_; }

private:
    gra_ctx *c_ = nullptr;
};

class Fwd {
public:
    // prisms: Nm x {x0,x1,y0,y1,z0,z1}; points: N_obs x {x,y,z}; field_flag
    // as GS::ComputationOptions.
    Fwd(Context &ctx, std::vector<double> prisms, std::vector<double> points,
        unsigned long long field_flag = GRA_GZ)
        : ctx_(ctx), prisms_(std::move(prisms)), points_(std::move(points)),
          flag_((uint32_t)field_flag)
    {
        Nm = (int64_t)prisms_.size() / 6;
        N_obs = (int64_t)points_.size() / 3;
        int nc = 0;
        for (int b = 1; b < 10; ++b) nc += (flag_ >> b) & 1u;
        Nd = nc * N_obs;
    }
    virtual ~Fwd() = default;

    void compute_G(gra_dtype store = GRA_F64)
    {
        check(ctx_.get(), gra_assemble(ctx_.get(), prisms_.data(), Nm, points_.data(), N_obs, flag_,
                                       store),
              "compute_G");
    }
    // compute_G by per-prism corner sums (gra_assemble_corners): any mesh,
    // points strictly above every prism.
    void compute_G_corners(gra_dtype store = GRA_F64)
    {
        check(ctx_.get(), gra_assemble_corners(ctx_.get(), prisms_.data(), Nm, points_.data(),
                                               N_obs, flag_, store),
              "compute_G_corners");
    }
    // compute_G for a tensor-product mesh whose prisms are the cells of the
    // node coordinates xs, ys, zs in Mesh::generate_regular_mesh order (the
    // Fwd's prisms are not read): corner sums shared through LDS
    // (gra_assemble_rectilinear).
    template <class V>
    void compute_G_rectilinear(const V &xs, const V &ys, const V &zs, gra_dtype store = GRA_F64)
    {
        const int64_t nx = (int64_t)xs.size() - 1, ny = (int64_t)ys.size() - 1,
                      nz = (int64_t)zs.size() - 1;
        if (nx * ny * nz != Nm) throw Error(GRA_BAD_ARG, "compute_G_rectilinear: cell count");
        check(ctx_.get(), gra_assemble_rectilinear(ctx_.get(), xs.data(), nx, ys.data(), ny,
                                                   zs.data(), nz, points_.data(), N_obs, flag_,
                                                   store),
              "compute_G_rectilinear");
    }
    template <class V> void G_vec_mul(const V &vec, V &product) const
    {
        if ((int64_t)vec.size() != Nm) throw Error(GRA_BAD_ARG, "G_vec_mul: size");
        product.resize(Nd);
        check(ctx_.get(), gra_G_vec_mul(ctx_.get(), vec.data(), product.data()), "G_vec_mul");
    }
    template <class V> void GT_vec_mul(const V &vec, V &product) const
    {
        if ((int64_t)vec.size() != Nd) throw Error(GRA_BAD_ARG, "GT_vec_mul: size");
        product.resize(Nm);
        check(ctx_.get(), gra_GT_vec_mul(ctx_.get(), vec.data(), product.data()), "GT_vec_mul");
    }
    template <class V> void col_norms(V &norms) const
    {
        norms.resize(Nm);
        check(ctx_.get(), gra_col_norms(ctx_.get(), norms.data()), "col_norms");
    }
    template <class V> V compute_gobs(const V &rho) const
    {
        V d;
        G_vec_mul(rho, d);
        return d;
    }
    unsigned int get_n_fields() const { return (unsigned)(Nd / (N_obs ? N_obs : 1)); }

    int64_t Nm = 0, N_obs = 0, Nd = 0;

protected:
    Context &ctx_;
    std::vector<double> prisms_, points_;
    uint32_t flag_;
};

class GaussNewtonInversion : public Fwd {
public:
    using Fwd::Fwd;

    bool use_cross_gradient_constraint = false;
    double a_crg = 1.0;
    std::vector<double> S_crg;   // empty = identity
    bool quirk_mode = true;      // reproduce the reference's operator exactly

    // Same argument list as GaussNewtonInversion::solve_cg; W* / T* / diagonal
    // operators are any compressed row-major sparse matrices (SMatrix).
    template <class V, class SM>
    V solve_cg(const double tol, const int &maxit, double &err, int &iterations,
               const double &lambda, const V &mk, const V &mref, const SM &Wd, const V &dobs,
               const SM &P, const SM &Ws, const SM &Wx, const SM &Wy, const SM &Wz,
               const SM &Tx0, const SM &Ty0, const SM &Tz0, const SM &WL_s, const SM &WL_x,
               const SM &WL_y, const SM &WL_z, const SM &WL_tx, const SM &WL_ty, const SM &WL_tz)
    {
        Csr ws = to_csr(Ws), wx = to_csr(Wx), wy = to_csr(Wy), wz = to_csr(Wz);
        gra_csr v[4] = {ws.view(), wx.view(), wy.view(), wz.view()};
        Csr tx, ty, tz;
        gra_csr tv[3];
        const gra_csr *tp[3] = {nullptr, nullptr, nullptr};
        if (use_cross_gradient_constraint) {
            tx = to_csr(Tx0);
            ty = to_csr(Ty0);
            tz = to_csr(Tz0);
            tv[0] = tx.view();
            tv[1] = ty.view();
            tv[2] = tz.view();
            tp[0] = &tv[0];
            tp[1] = &tv[1];
            tp[2] = &tv[2];
        }
        check(ctx_.get(),
              gra_set_regularization(ctx_.get(), &v[0], &v[1], &v[2], &v[3], tp[0], tp[1], tp[2]),
              "set_regularization");
        std::vector<double> wd = diagonal(Wd), p = diagonal(P), ls = diagonal(WL_s),
                            lx = diagonal(WL_x), ly = diagonal(WL_y), lz = diagonal(WL_z),
                            ltx = diagonal(WL_tx), lty = diagonal(WL_ty), ltz = diagonal(WL_tz);
        gra_cg_args a{};
        a.tol = tol;
        a.maxit = maxit;
        a.lambda = lambda;
        a.mk = mk.data();
        a.mref = mref.data();
        a.Wd = wd.data();
        a.dobs = dobs.data();
        a.P = p.data();
        a.WL_s = ls.data(); a.WL_x = lx.data(); a.WL_y = ly.data(); a.WL_z = lz.data();
        a.WL_tx = ltx.data(); a.WL_ty = lty.data(); a.WL_tz = ltz.data();
        a.use_cross_gradient = use_cross_gradient_constraint ? 1 : 0;
        a.a_crg = a_crg;
        a.S_crg = S_crg.empty() ? nullptr : S_crg.data();
        a.quirk_mode = quirk_mode ? 1 : 0;
        V dx;
        dx.resize(Nm);
        gra_status s = gra_solve_cg(ctx_.get(), &a, dx.data(), &err, &iterations);
        if (s != GRA_NUMERIC_BREAKDOWN) check(ctx_.get(), s, "solve_cg");
        return dx;
    }
};

}  // namespace graeng
