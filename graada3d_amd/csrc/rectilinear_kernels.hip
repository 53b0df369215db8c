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

__global__ void __launch_bounds__(kBlk)
node_forward_kernel(const double *__restrict__ xs, const double *__restrict__ ys,
                    const double *__restrict__ zs, int64_t nx, int64_t ny, int64_t nz,
                    const double *__restrict__ w, const double *__restrict__ pts, int64_t N_obs,
                    uint32_t flag, double *__restrict__ out)
{
    __shared__ double zl[kMaxLevels];
    __shared__ double lg_G[kLogTable], lg_T[kLogTable];
    const int64_t ip = blockIdx.x;
    if (ip >= N_obs) return;
    for (int t = threadIdx.x; t < kLogTable; t += kBlk) {
        lg_G[t] = kLogG[t];
        lg_T[t] = kLogT[t];
    }
    const double px = pts[3 * ip], py = pts[3 * ip + 1], pz = pts[3 * ip + 2];
    const int NY = (int)ny + 1, NZ = (int)nz + 1, NCOL = ((int)nx + 1) * NY;
    const bool lds_levels = NZ <= kMaxLevels;
    if (lds_levels) {
        for (int k = threadIdx.x; k < NZ; k += kBlk) zl[k] = zs[k] - pz;
    }
    __syncthreads();
    // which transcendental terms the requested components need (uniform)
    const bool need_lx = flag & ((1u << 1) | (1u << 3) | (1u << 6));
    const bool need_ly = flag & ((1u << 1) | (1u << 2) | (1u << 5));
    const bool need_lz = flag & ((1u << 2) | (1u << 3) | (1u << 8));
    const bool need_axy = flag & ((1u << 1) | (1u << 4));
    const bool need_ayz = flag & ((1u << 2) | (1u << 7));
    const bool need_axz = flag & ((1u << 3) | (1u << 9));
    const bool all_atan = need_axy && need_ayz && need_axz;
    const bool any_atan = need_axy || need_ayz || need_axz;
    double acc[10];
#pragma unroll
    for (int q = 0; q < 10; ++q) acc[q] = 0.0;
    for (int col = threadIdx.x; col < NCOL; col += kBlk) {
        const int i = col / NY, j = col - i * NY;
        const double x = xs[i] - px, y = ys[j] - py;
        const double xy2 = fma(x, x, y * y);
        const bool axis = (x == 0.0) && (y == 0.0);
        const double xy = x * y;
        // (pi/2) sgn(x) sgn(y) with the sign of a zero: the limits 1/(+-0)
        // gives the atans it replaces
        const double half_pi_s = copysign(1.5707963267948966, x) * copysign(1.0, y);
        for (int k = 0; k < NZ; ++k) {
            const double wn = w[(int64_t)k * NCOL + col];
            const double z = lds_levels ? zl[k] : zs[k] - pz;
            const double r = sqrt(xy2 + z * z);
            const double lx = need_lx ? table_log(x + r, lg_G, lg_T) : 0.0;
            const double ly = need_ly ? table_log(y + r, lg_G, lg_T) : 0.0;
            const double lz = need_lz ? table_log(z + r, lg_G, lg_T) : 0.0;
            double axy = 0.0, ayz = 0.0, axz = 0.0;
            if (any_atan) {
                if (need_axy) axy = atan_ratio(xy, z * r);
                if (axis) {
                    const double a = atan_ratio(z, r);
                    ayz = a;
                    axz = a;
                } else if (all_atan) {
                    ayz = atan_ratio(y * z, x * r);
                    axz = half_pi_s - axy - ayz;
                } else {
                    if (need_ayz) ayz = atan_ratio(y * z, x * r);
                    if (need_axz) axz = atan_ratio(x * z, y * r);
                }
            }
            if (flag & (1u << 1)) acc[1] += wn * (x * ly + y * lx - z * axy);
            if (flag & (1u << 2)) acc[2] += wn * (y * lz + z * ly - x * ayz);
            if (flag & (1u << 3)) acc[3] += wn * (z * lx + x * lz - y * axz);
            if (flag & (1u << 4)) acc[4] += wn * axy;
            if (flag & (1u << 5)) acc[5] += wn * ly;
            if (flag & (1u << 6)) acc[6] += wn * lx;
            if (flag & (1u << 7)) acc[7] += wn * ayz;
            if (flag & (1u << 8)) acc[8] += wn * lz;
            if (flag & (1u << 9)) acc[9] += wn * axz;
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
        const double cg = -kG0 * 1e5, ct = kG0 * 1e9;
        const double scale[10] = {0.0, cg, cg, cg, -ct, ct, ct, -ct, ct, -ct};
        int nc = 0;
        for (int bit = 1; bit < 10; ++bit) nc += (flag >> bit) & 1u;
        int col = 0;
        for (int bit = 1; bit < 10; ++bit) {
            if (!(flag & (1u << bit))) continue;
            double s = red[0][bit];
            for (int q = 1; q < kBlk / 64; ++q) s += red[q][bit];
            out[ip * nc + col] = scale[bit] * s;
            ++col;
        }
    }
}

}  // namespace

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
