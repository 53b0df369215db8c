// Matrix-free forward for rectilinear (tensor-product) meshes by summing the
// prism corner function over shared mesh nodes.
//
// A prism's field is the signed sum over its 8 corners of a closed-form
// corner function F (the same quantity the reference integrates over faces and
// edges, GravFormula.cpp / Integral.cpp).  On a mesh whose cells are the
// tensor product of 1-D partitions xs, ys, zs every interior node is a corner
// of 8 cells, so
//
//     d_k(p) = c_k * sum_nodes w_n * F_k(node - p),
//     w = Dx Dy Dz rho,  (D f)[i] = f[i-1] - f[i] (f = 0 outside the mesh),
//
// one corner evaluation per node instead of 8 per cell.  Corner functions
// (x, y, z = node - p, r = |(x, y, z)|, z positive down):
//   gz: x ln(y+r) + y ln(x+r) - z atan(xy/(zr))     gx, gy: cyclic
//   Tzz: atan(xy/(zr))   Txx: atan(yz/(xr))   Tyy: atan(xz/(yr))
//   Txy: ln(z+r)         Txz: ln(y+r)         Tyz: ln(x+r)
// with c = -G0*1e5 for g (mGal) and -G0*1e9 (Tzz, Txx, Tyy) / +G0*1e9 (Txy,
// Txz, Tyz) for T (Eotvos), matching the reference's units and signs.
//
// Points must lie strictly above the mesh (z_p < zs[0]): then z > 0 and r > 0
// for every node and only x = 0 or y = 0 (the point over a node plane) needs
// care.  IEEE gives x = +0 (a - a), i.e. the limit from the point shifted by
// -0 in x; where x = y = 0 both ratio terms take the limit along the shift
// (-d, -d, 0), atan(z/r) -> pi/4.  The field is continuous outside the body,
// so one consistent limit for all nodes yields the field at p.
//
// The same corner sums assemble G (node_assemble_kernel): entry (k*N_obs + i,
// cell) = c_k * sum over the cell's 8 corners of s * F_k(corner - p_i), s = +1
// at the corner (x1, y1, z1) and alternating, with the corner values of a
// tile of cells computed once per point and shared through LDS.
#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include "engine.hpp"
#include "fastatan.hpp"
#include "fastlog.hpp"

namespace gra {

namespace {

constexpr int kBlk = 256;
constexpr double kG0 = 6.673e-11;

// The log of the corner terms: table_log (fastlog.hpp) with the table
// copied to LDS per block.
constexpr int kLogTable = kLogTableN;

// Node weights, stored level by level: w[k * (NX*NY) + i * NY + j], so the
// threads of a block (consecutive node columns) read consecutive addresses.
__global__ void node_weights_kernel(const double *__restrict__ rho, int64_t nx, int64_t ny,
                                    int64_t nz, double *__restrict__ w)
{
    const int64_t NX = nx + 1, NY = ny + 1, NZ = nz + 1;
    int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= NX * NY * NZ) return;
    const int64_t col = n % (NX * NY), k = n / (NX * NY);
    const int64_t i = col / NY, j = col % NY;
    // w = sum over the 8 cells (i-di, j-dj, k-dk), di,dj,dk in {0,1}, of
    // s(di) s(dj) s(dk) rho, s(0) = -1, s(1) = +1
    double acc = 0.0;
#pragma unroll
    for (int di = 0; di < 2; ++di)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
            for (int dk = 0; dk < 2; ++dk) {
                const int64_t ci = i - di, cj = j - dj, ck = k - dk;
                if (ci < 0 || ci >= nx || cj < 0 || cj >= ny || ck < 0 || ck >= nz) continue;
                const double s = ((di + dj + dk) & 1) ? 1.0 : -1.0;
                acc += s * rho[(ci * ny + cj) * nz + ck];
            }
    w[n] = acc;
}

// One block per observation point; a thread takes node columns (i, j), with
// x and y fixed, and walks their nz+1 nodes.  The atans are atan_ratio
// (fastatan.hpp: one division each, no library call) and the logs table_log
// (fastlog.hpp, table in LDS); when all three atans are needed the third
// follows from the corner identity
//   atan(xy/(zr)) + atan(yz/(xr)) + atan(xz/(yr)) = (pi/2) sgn(x) sgn(y)
// (z > 0; sgn(+0) = +1, which is also the limit the x = 0 / y = 0 / axis
// cases take).
constexpr int kMaxLevels = 1024;

// Which transcendental terms the requested components need.
struct NodeTerms {
    bool lx, ly, lz, axy, ayz, axz, all_atan, any_atan;
    __device__ explicit NodeTerms(uint32_t flag)
        : lx(flag & ((1u << 1) | (1u << 3) | (1u << 6))),
          ly(flag & ((1u << 1) | (1u << 2) | (1u << 5))),
          lz(flag & ((1u << 2) | (1u << 3) | (1u << 8))),
          axy(flag & ((1u << 1) | (1u << 4))),
          ayz(flag & ((1u << 2) | (1u << 7))),
          axz(flag & ((1u << 3) | (1u << 9)))
    {
        all_atan = axy && ayz && axz;
        any_atan = axy || ayz || axz;
    }
};

// The transcendental terms at one node: x, y, z = node - point (z > 0);
// xy2 = x^2 + y^2 (fma), axis = (x == 0 && y == 0), half_pi_s = (pi/2) sgn(x)
// sgn(y) with signed zeros.
struct CornerTerms {
    double lx, ly, lz, axy, ayz, axz;
};

__device__ inline CornerTerms corner_terms(double x, double y, double z, double xy2, bool axis,
                                          double half_pi_s, const NodeTerms &nt,
                                          const double *lg_G, const double *lg_T)
{
    CornerTerms c;
    const double r = sqrt(xy2 + z * z);
    // ln(u + r) for u < 0 would lose digits to the cancellation u + r (a
    // point far along -u): there it is ln((r^2 - u^2) / (r - u)), the
    // numerator formed from the other two squares (positive: z > 0).
    // z > 0 needs no care.
    if (nt.lx) c.lx = table_log(x >= 0.0 ? x + r : fma(y, y, z * z) / (r - x), lg_G, lg_T);
    else c.lx = 0.0;
    if (nt.ly) c.ly = table_log(y >= 0.0 ? y + r : fma(x, x, z * z) / (r - y), lg_G, lg_T);
    else c.ly = 0.0;
    c.lz = nt.lz ? table_log(z + r, lg_G, lg_T) : 0.0;
    c.axy = c.ayz = c.axz = 0.0;
    if (nt.any_atan) {
        if (nt.axy) c.axy = atan_ratio(x * y, z * r);
        if (axis) {
            const double a = atan_ratio(z, r);
            c.ayz = a;
            c.axz = a;
        } else if (nt.all_atan) {
            c.ayz = atan_ratio(y * z, x * r);
            c.axz = half_pi_s - c.axy - c.ayz;
        } else {
            if (nt.ayz) c.ayz = atan_ratio(y * z, x * r);
            if (nt.axz) c.axz = atan_ratio(x * z, y * r);
        }
    }
    return c;
}

// Corner function of component `bit` from the terms.
__device__ inline double corner_component(int bit, double x, double y, double z,
                                          const CornerTerms &c)
{
    switch (bit) {
    case 1: return x * c.ly + y * c.lx - z * c.axy;
    case 2: return y * c.lz + z * c.ly - x * c.ayz;
    case 3: return z * c.lx + x * c.lz - y * c.axz;
    case 4: return c.axy;
    case 5: return c.ly;
    case 6: return c.lx;
    case 7: return c.ayz;
    case 8: return c.lz;
    default: return c.axz;
    }
}

// Output scale of component bit: -G0*1e5 (g, mGal), -/+G0*1e9 (T, Eotvos).
__device__ inline double comp_scale(int bit)
{
    const double cg = -kG0 * 1e5, ct = kG0 * 1e9;
    switch (bit) {
    case 1: case 2: case 3: return cg;
    case 4: case 7: case 9: return -ct;
    default: return ct;
    }
}

// The forward sum of one observation point (one workgroup): every mesh node's
// weight times the corner function, summed over the nodes column by column
// by the workgroup's threads, then reduced in a fixed order.  Block-uniform
// arguments (it synchronises the workgroup).
__device__ inline void node_forward_point(const double *__restrict__ xs,
                                          const double *__restrict__ ys,
                                          const double *__restrict__ zs, int64_t nx, int64_t ny,
                                          int64_t nz, const double *__restrict__ w, double px,
                                          double py, double pz, uint32_t flag,
                                          double *__restrict__ out_row)
{
    __shared__ double zl[kMaxLevels];
    __shared__ double lg_G[kLogTable], lg_T[kLogTable];
    for (int t = threadIdx.x; t < kLogTable; t += kBlk) {
        lg_G[t] = kLogG[t];
        lg_T[t] = kLogT[t];
    }
    const int NY = (int)ny + 1, NZ = (int)nz + 1, NCOL = ((int)nx + 1) * NY;
    const bool lds_levels = NZ <= kMaxLevels;
    if (lds_levels) {
        for (int k = threadIdx.x; k < NZ; k += kBlk) zl[k] = zs[k] - pz;
    }
    __syncthreads();
    const NodeTerms nt(flag);   // uniform
    double acc[10];
#pragma unroll
    for (int q = 0; q < 10; ++q) acc[q] = 0.0;
    for (int col = threadIdx.x; col < NCOL; col += kBlk) {
        const int i = col / NY, j = col - i * NY;
        const double x = xs[i] - px, y = ys[j] - py;
        const double xy2 = fma(x, x, y * y);
        const bool axis = (x == 0.0) && (y == 0.0);
        // (pi/2) sgn(x) sgn(y) with the sign of a zero: the limits 1/(+-0)
        // gives the atans it replaces
        const double half_pi_s = copysign(1.5707963267948966, x) * copysign(1.0, y);
        for (int k = 0; k < NZ; ++k) {
            const double wn = w[(int64_t)k * NCOL + col];
            const double z = lds_levels ? zl[k] : zs[k] - pz;
            const CornerTerms ct = corner_terms(x, y, z, xy2, axis, half_pi_s, nt, lg_G, lg_T);
#pragma unroll
            for (int bit = 1; bit < 10; ++bit)
                if (flag & (1u << bit)) acc[bit] += wn * corner_component(bit, x, y, z, ct);
        }
    }
    __shared__ double red[kBlk / 64][10];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int bit = 1; bit < 10; ++bit) {
        if (!(flag & (1u << bit))) continue;
        double v = acc[bit];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
        if (lane == 0) red[wave][bit] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int col = 0;
        for (int bit = 1; bit < 10; ++bit) {
            if (!(flag & (1u << bit))) continue;
            double s = red[0][bit];
            for (int q = 1; q < kBlk / 64; ++q) s += red[q][bit];
            out_row[col] = comp_scale(bit) * s;
            ++col;
        }
    }
}

__global__ void __launch_bounds__(kBlk)
node_forward_kernel(const double *__restrict__ xs, const double *__restrict__ ys,
                    const double *__restrict__ zs, int64_t nx, int64_t ny, int64_t nz,
                    const double *__restrict__ w, const double *__restrict__ pts, int64_t N_obs,
                    uint32_t flag, double *__restrict__ out)
{
    const int64_t ip = blockIdx.x;
    if (ip >= N_obs) return;
    int nc = 0;
    for (int bit = 1; bit < 10; ++bit) nc += (flag >> bit) & 1u;
    node_forward_point(xs, ys, zs, nx, ny, nz, w, pts[3 * ip], pts[3 * ip + 1], pts[3 * ip + 2],
                       flag, out + ip * nc);
}

// Several forward requests in one launch: workgroup b takes point
// b - block0 of the item whose block range holds b (items sorted by block0).
__global__ void __launch_bounds__(kBlk)
node_forward_batch_kernel(const FwdItem *__restrict__ items, int n_items)
{
    int lo = 0, hi = n_items - 1;
    const int64_t b = blockIdx.x;
    while (lo < hi) {                       // last item with block0 <= b
        const int mid = (lo + hi + 1) / 2;
        if (items[mid].block0 <= b) lo = mid;
        else hi = mid - 1;
    }
    const FwdItem &it = items[lo];
    const int64_t ip = b - it.block0;
    if (ip >= it.N_obs) return;
    int nc = 0;
    for (int bit = 1; bit < 10; ++bit) nc += (it.flag >> bit) & 1u;
    node_forward_point(it.xs, it.ys, it.zs, it.nx, it.ny, it.nz, it.w, it.pts[3 * ip],
                       it.pts[3 * ip + 1], it.pts[3 * ip + 2], it.flag, it.out + ip * nc);
}

// (a, b, k) of the flat index i = (a * nb + b) * nk + k, advanced by kBlk per
// step with carries (the step's own (da, db, dk) computed once).
struct Walk3 {
    int a, b, k, da, db, dk, nb, nk;
    __device__ Walk3(int i, int nb_, int nk_) : nb(nb_), nk(nk_)
    {
        k = i % nk;
        b = (i / nk) % nb;
        a = i / (nk * nb);
        dk = kBlk % nk;
        db = (kBlk / nk) % nb;
        da = kBlk / (nk * nb);
    }
    __device__ inline void next()
    {
        k += dk;
        int carry = k >= nk;
        k -= carry ? nk : 0;
        b += db + carry;
        carry = b >= nb;
        b -= carry ? nb : 0;
        a += da + carry;
    }
};

template <typename T> __device__ inline T store_as(double v);
template <> __device__ inline double store_as<double>(double v) { return v; }
template <> __device__ inline float store_as<float>(double v) { return (float)v; }
template <> __device__ inline __hip_bfloat16 store_as<__hip_bfloat16>(double v)
{
    return __float2bfloat16((float)v);
}

// Two neighbouring G entries (g even-aligned) in one store.
template <typename T> __device__ inline void store_pair(T *g, T a, T b)
{
    if constexpr (sizeof(T) == 8) {
        typedef double P __attribute__((ext_vector_type(2)));
        *reinterpret_cast<P *>(g) = P{(double)a, (double)b};
    } else if constexpr (sizeof(T) == 4) {
        typedef float P __attribute__((ext_vector_type(2)));
        *reinterpret_cast<P *>(g) = P{(float)a, (float)b};
    } else {
        const uint32_t lo = __builtin_bit_cast(unsigned short, a);
        const uint32_t hi = __builtin_bit_cast(unsigned short, b);
        *reinterpret_cast<uint32_t *>(g) = lo | (hi << 16);
    }
}

// G by corner sums.  One workgroup per (point, tile of tile x tile cell
// columns with all nz levels): the corner values of the tile's
// (tile+1)^2 (nz+1) nodes go to LDS once (component-major, nmax per
// component), then each cell's entry is the signed 8-corner sum, formed as
// differences along x, then y, then z (differences of neighbouring values
// first).  Threads take consecutive cells of one x column of the tile, i.e.
// consecutive G columns.
template <typename T>
__global__ void __launch_bounds__(kBlk)
node_assemble_kernel(const double *__restrict__ xs, const double *__restrict__ ys,
                     const double *__restrict__ zs, int64_t nx, int64_t ny, int64_t nz,
                     const double *__restrict__ pts, int64_t N_obs, int64_t p0, uint32_t flag,
                     int tile, int64_t tiles_y, int nmax, T *__restrict__ G, int64_t ld)
{
    extern __shared__ double vals[];   // [nc][nmax]
    __shared__ double lg_G[kLogTable], lg_T[kLogTable];
    const int64_t ip = p0 + blockIdx.y, t = blockIdx.x;
    if (ip >= N_obs) return;
    for (int q = threadIdx.x; q < kLogTable; q += kBlk) {
        lg_G[q] = kLogG[q];
        lg_T[q] = kLogT[q];
    }
    const int64_t i0 = (t / tiles_y) * tile, j0 = (t % tiles_y) * tile;
    const int cx = (int)(nx - i0 < tile ? nx - i0 : tile);
    const int cy = (int)(ny - j0 < tile ? ny - j0 : tile);
    const int NZ = (int)nz + 1, plane = (cy + 1) * NZ, nnodes = (cx + 1) * plane;
    const double px = pts[3 * ip], py = pts[3 * ip + 1], pz = pts[3 * ip + 2];
    const NodeTerms nt(flag);
    __syncthreads();
    // node n = (a * (cy+1) + b) * NZ + k; (a, b, k) advanced by kBlk per step
    // without integer divisions in the loop
    Walk3 wn(threadIdx.x, cy + 1, NZ);
    for (int n = threadIdx.x; n < nnodes; n += kBlk, wn.next()) {
        const int a = wn.a, b = wn.b, k = wn.k;
        const double x = xs[i0 + a] - px, y = ys[j0 + b] - py, z = zs[k] - pz;
        const double xy2 = fma(x, x, y * y);
        const bool axis = (x == 0.0) && (y == 0.0);
        const double half_pi_s = copysign(1.5707963267948966, x) * copysign(1.0, y);
        const CornerTerms ct = corner_terms(x, y, z, xy2, axis, half_pi_s, nt, lg_G, lg_T);
        int q = 0;
#pragma unroll
        for (int bit = 1; bit < 10; ++bit)
            if (flag & (1u << bit)) vals[q++ * nmax + n] = corner_component(bit, x, y, z, ct);
    }
    __syncthreads();
    // A thread takes the cells k and k+1 of a column: with
    // e(l) = (x-difference at y1) - (x-difference at y0) of level l, a cell's
    // sum is e(k+1) - e(k), so the shared level is read once (12 LDS reads
    // for two cells instead of 16; the same operations as cell by cell).
    const int nzc = (int)nz, npair = (nzc + 1) >> 1, nwork = cx * cy * npair;
    Walk3 wc(threadIdx.x, cy, npair);
    for (int c = threadIdx.x; c < nwork; c += kBlk, wc.next()) {
        const int a = wc.a, b = wc.b, k = 2 * wc.k;
        const bool second = k + 1 < nzc;          // level k + 2 <= nz exists
        const int n0 = a * plane + b * NZ + k;
        const int64_t j = ((i0 + a) * ny + (j0 + b)) * nz + k;
        int q = 0;
#pragma unroll
        for (int bit = 1; bit < 10; ++bit) {
            if (!(flag & (1u << bit))) continue;
            const double *v = vals + q * nmax + n0;
            const double e0 = (v[plane + NZ] - v[NZ]) - (v[plane] - v[0]);
            const double e1 = (v[plane + NZ + 1] - v[NZ + 1]) - (v[plane + 1] - v[1]);
            T *g = G + ((int64_t)q * N_obs + ip) * ld + j;
            const T y0 = store_as<T>(comp_scale(bit) * (e1 - e0));
            if (second) {
                const double e2 = (v[plane + NZ + 2] - v[NZ + 2]) - (v[plane + 2] - v[2]);
                const T y1 = store_as<T>(comp_scale(bit) * (e2 - e1));
                if ((j & 1) == 0) {
                    store_pair(g, y0, y1);        // G rows start at even columns
                } else {
                    g[0] = y0;
                    g[1] = y1;
                }
            } else {
                g[0] = y0;
            }
            ++q;
        }
    }
}

// Corner sums for any prisms (cell_corner_kernel): a thread takes one cell
// for one point and evaluates the cell's eight corners itself (no sharing
// between cells: octree and other non-tensor meshes), combined in the
// difference order of node_assemble_kernel -- x differences first, then y,
// then z -- so a cell of a tensor-product mesh gets the same value both ways.
// Threads take consecutive cells of one point: consecutive G columns.
// sum[bit - 1] receives the unscaled corner sum of each requested component.
__device__ inline void prism_corner_sums(double px, double py, double pz,
                                         const double *__restrict__ pr, uint32_t flag,
                                         const double *lg_G, const double *lg_T, double sum[9])
{
    const double xv[2] = {pr[0] - px, pr[1] - px};
    const double yv[2] = {pr[2] - py, pr[3] - py};
    const double zv[2] = {pr[4] - pz, pr[5] - pz};
    const NodeTerms nt(flag);
    double e0[9];
#pragma unroll
    for (int kk = 0; kk < 2; ++kk) {
#pragma unroll
        for (int bb = 0; bb < 2; ++bb) {
            // x difference F(x1, y_bb, z_kk) - F(x0, y_bb, z_kk) per component
            double d[9];
#pragma unroll
            for (int aa = 0; aa < 2; ++aa) {
                const double x = xv[aa], y = yv[bb], z = zv[kk];
                const double xy2 = fma(x, x, y * y);
                const bool axis = (x == 0.0) && (y == 0.0);
                const double half_pi_s = copysign(1.5707963267948966, x) * copysign(1.0, y);
                const CornerTerms ct = corner_terms(x, y, z, xy2, axis, half_pi_s, nt, lg_G, lg_T);
#pragma unroll
                for (int bit = 1; bit < 10; ++bit) {
                    if (!(flag & (1u << bit))) continue;
                    const double v = corner_component(bit, x, y, z, ct);
                    d[bit - 1] = aa ? v - d[bit - 1] : v;
                }
            }
#pragma unroll
            for (int bit = 1; bit < 10; ++bit) {
                if (!(flag & (1u << bit))) continue;
                if (bb == 0) sum[bit - 1] = d[bit - 1];
                else if (kk == 0) e0[bit - 1] = d[bit - 1] - sum[bit - 1];
                else sum[bit - 1] = (d[bit - 1] - sum[bit - 1]) - e0[bit - 1];
            }
        }
    }
}

template <typename T>
__global__ void __launch_bounds__(kBlk)
cell_corner_kernel(const double *__restrict__ prisms, int64_t Nm, const double *__restrict__ pts,
                   int64_t N_obs, int64_t p0, uint32_t flag, T *__restrict__ G, int64_t ld)
{
    __shared__ double lg_G[kLogTable], lg_T[kLogTable];
    for (int q = threadIdx.x; q < kLogTable; q += kBlk) {
        lg_G[q] = kLogG[q];
        lg_T[q] = kLogT[q];
    }
    __syncthreads();
    const int64_t ip = p0 + blockIdx.y;
    const int64_t j = (int64_t)blockIdx.x * kBlk + threadIdx.x;
    if (ip >= N_obs || j >= Nm) return;
    double sum[9];
    prism_corner_sums(pts[3 * ip], pts[3 * ip + 1], pts[3 * ip + 2], prisms + 6 * j, flag, lg_G,
                      lg_T, sum);
    int q = 0;
#pragma unroll
    for (int bit = 1; bit < 10; ++bit) {
        if (!(flag & (1u << bit))) continue;
        G[((int64_t)q * N_obs + ip) * ld + j] = store_as<T>(comp_scale(bit) * sum[bit - 1]);
        ++q;
    }
}

// The batched form over uneven instances: blocks are distributed over the
// items as in the batched face/edge assembly (item = last block0 <= block;
// within an item, blocks walk the point rows, ceil(Nm / 256) per row).
template <typename T>
__global__ void __launch_bounds__(kBlk)
cell_corner_batch_kernel(const BatchItem *__restrict__ items, int n_items)
{
    __shared__ double lg_G[kLogTable], lg_T[kLogTable];
    __shared__ int s_item;
    for (int q = threadIdx.x; q < kLogTable; q += kBlk) {
        lg_G[q] = kLogG[q];
        lg_T[q] = kLogT[q];
    }
    const int64_t b = blockIdx.x;
    if (threadIdx.x == 0) {
        int lo = 0, hi = n_items - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (items[mid].block0 <= b) lo = mid;
            else hi = mid - 1;
        }
        s_item = lo;
    }
    __syncthreads();
    const BatchItem it = items[s_item];
    const int64_t lb = b - it.block0;
    const int64_t i = lb / it.bpr;
    const int64_t j = (lb - i * it.bpr) * kBlk + threadIdx.x;
    if (i >= it.N_obs || j >= it.Nm) return;
    double sum[9];
    prism_corner_sums(it.pts[3 * i], it.pts[3 * i + 1], it.pts[3 * i + 2], it.prisms + 6 * j,
                      it.flag, lg_G, lg_T, sum);
    T *G = static_cast<T *>(it.G);
    int q = 0;
#pragma unroll
    for (int bit = 1; bit < 10; ++bit) {
        if (!(it.flag & (1u << bit))) continue;
        G[((int64_t)q * it.N_obs + i) * it.ld + j] = store_as<T>(comp_scale(bit) * sum[bit - 1]);
        ++q;
    }
}

}  // namespace

// Tile side (cell columns, <= 16): the largest whose corner values fit 60 KB
// of LDS and whose node phase takes at most 8 passes of the workgroup (2,048
// nodes); without such a tile (deep meshes), the largest that fits.  Swept
// on MI355X: gz+Tzz on the 40x40x20 mesh peaks at 8 (1.75e11 entries/s; 10:
// 1.72e11, 12: 1.01e11, 5: 1.36e11), nine components on 64x64x32 at 4
// (2.63e11; 2: 2.01e11).  0 if not even one column fits.
int node_assemble_tile(int ncomp, int64_t nz)
{
    constexpr double kLdsBudget = 61440.0;
    constexpr double kMaxNodes = 2048.0;
    int fit = 0;
    for (int t = 16; t >= 1; --t) {
        const double nodes = (double)(t + 1) * (t + 1) * (double)(nz + 1);
        if (8.0 * ncomp * nodes > kLdsBudget) continue;
        if (!fit) fit = t;
        if (nodes <= kMaxNodes) return t;
    }
    return fit;
}

hipError_t launch_node_assemble(const double *xs, const double *ys, const double *zs, int64_t nx,
                                int64_t ny, int64_t nz, const double *pts, int64_t N_obs,
                                uint32_t flag, gra_dtype store, void *G, int64_t ld, int tile,
                                hipStream_t st)
{
    int nc = 0;
    for (int bit = 1; bit < 10; ++bit) nc += (flag >> bit) & 1u;
    const int nmax = (tile + 1) * (tile + 1) * (int)(nz + 1);
    const size_t lds = sizeof(double) * (size_t)nc * (size_t)nmax;
    const int64_t tiles_x = (nx + tile - 1) / tile, tiles_y = (ny + tile - 1) / tile;
    const int64_t ntiles = tiles_x * tiles_y;
    constexpr int64_t kMaxY = 65535;   // points per launch (grid y)
    for (int64_t p0 = 0; p0 < N_obs; p0 += kMaxY) {
        const int64_t np = N_obs - p0 < kMaxY ? N_obs - p0 : kMaxY;
        const dim3 grid((unsigned)ntiles, (unsigned)np), block(kBlk);
        switch (store) {
        case GRA_F64:
            hipLaunchKernelGGL(node_assemble_kernel<double>, grid, block, lds, st, xs, ys, zs, nx,
                               ny, nz, pts, N_obs, p0, flag, tile, tiles_y, nmax, (double *)G, ld);
            break;
        case GRA_F32:
            hipLaunchKernelGGL(node_assemble_kernel<float>, grid, block, lds, st, xs, ys, zs, nx,
                               ny, nz, pts, N_obs, p0, flag, tile, tiles_y, nmax, (float *)G, ld);
            break;
        default:
            hipLaunchKernelGGL(node_assemble_kernel<__hip_bfloat16>, grid, block, lds, st, xs, ys,
                               zs, nx, ny, nz, pts, N_obs, p0, flag, tile, tiles_y, nmax,
                               (__hip_bfloat16 *)G, ld);
            break;
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_cell_corner(const double *prisms, int64_t Nm, const double *pts, int64_t N_obs,
                              uint32_t flag, gra_dtype store, void *G, int64_t ld, hipStream_t st)
{
    constexpr int64_t kMaxY = 65535;   // points per launch (grid y)
    const unsigned bx = (unsigned)((Nm + kBlk - 1) / kBlk);
    for (int64_t p0 = 0; p0 < N_obs; p0 += kMaxY) {
        const int64_t np = N_obs - p0 < kMaxY ? N_obs - p0 : kMaxY;
        const dim3 grid(bx, (unsigned)np), block(kBlk);
        switch (store) {
        case GRA_F64:
            hipLaunchKernelGGL(cell_corner_kernel<double>, grid, block, 0, st, prisms, Nm, pts,
                               N_obs, p0, flag, (double *)G, ld);
            break;
        case GRA_F32:
            hipLaunchKernelGGL(cell_corner_kernel<float>, grid, block, 0, st, prisms, Nm, pts,
                               N_obs, p0, flag, (float *)G, ld);
            break;
        default:
            hipLaunchKernelGGL(cell_corner_kernel<__hip_bfloat16>, grid, block, 0, st, prisms, Nm,
                               pts, N_obs, p0, flag, (__hip_bfloat16 *)G, ld);
            break;
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_cell_corner_batch(const BatchItem *d_items, int n_items, int64_t total_blocks,
                                    gra_dtype store, hipStream_t st)
{
    if (total_blocks <= 0) return hipSuccess;
    const dim3 grid((unsigned)total_blocks), block(kBlk);
    switch (store) {
    case GRA_F64:
        hipLaunchKernelGGL(cell_corner_batch_kernel<double>, grid, block, 0, st, d_items, n_items);
        break;
    case GRA_F32:
        hipLaunchKernelGGL(cell_corner_batch_kernel<float>, grid, block, 0, st, d_items, n_items);
        break;
    default:
        hipLaunchKernelGGL(cell_corner_batch_kernel<__hip_bfloat16>, grid, block, 0, st, d_items,
                           n_items);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_node_forward_batch(const FwdItem *h_items, const FwdItem *d_items, int n_items,
                                     int64_t total_blocks, hipStream_t st)
{
    for (int q = 0; q < n_items; ++q) {
        const FwdItem &it = h_items[q];
        const int64_t NN = (it.nx + 1) * (it.ny + 1) * (it.nz + 1);
        hipLaunchKernelGGL(node_weights_kernel, dim3((unsigned)((NN + kBlk - 1) / kBlk)),
                           dim3(kBlk), 0, st, it.rho, it.nx, it.ny, it.nz, it.w);
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || total_blocks == 0) return e;
    hipLaunchKernelGGL(node_forward_batch_kernel, dim3((unsigned)total_blocks), dim3(kBlk), 0, st,
                       d_items, n_items);
    return hipGetLastError();
}

hipError_t launch_node_forward(const double *xs, const double *ys, const double *zs, int64_t nx,
                               int64_t ny, int64_t nz, const double *rho, double *w,
                               const double *pts, int64_t N_obs, uint32_t flag, double *out,
                               hipStream_t st)
{
    const int64_t NN = (nx + 1) * (ny + 1) * (nz + 1);
    hipLaunchKernelGGL(node_weights_kernel, dim3((unsigned)((NN + kBlk - 1) / kBlk)), dim3(kBlk),
                       0, st, rho, nx, ny, nz, w);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || N_obs == 0) return e;
    hipLaunchKernelGGL(node_forward_kernel, dim3((unsigned)N_obs), dim3(kBlk), 0, st, xs, ys, zs,
                       nx, ny, nz, w, pts, N_obs, flag, out);
    return hipGetLastError();
}

}  // namespace gra
