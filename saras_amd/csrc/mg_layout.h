// Red-black split layout of one multigrid level and the ghost rules shared by host and device.
//
// A level has nx x ny x nz core nodes (logical index 0..n-1) and one ghost layer on each side
// (logical -1 and n).  Padded index p = logical + 1 lies in [0, n+1].  Node (ip, jp, kp) has colour
//   c = (ip + jp + kp + 1) & 1          (c = 0 <=> logical i+j+k even: "red")
// and is stored at half-row position h = kp >> 1 of the colour-c half-row of column (ip, jp):
//   index = ((ip * (ny+2) + jp) * 2 + c) * khs + h
// A red-black half sweep of colour c therefore streams colour c of f and u once and colour 1-c of
// u once: 12 B (FP64) per node and half sweep, against 24 B for the same sweep over the natural
// interleaved layout.  The x/y neighbours of a node sit at the same h in the opposite-colour
// half-row of the neighbouring column; the z neighbours are two adjacent entries of the
// opposite-colour half-row of the same column.
#pragma once

#include <hip/hip_runtime.h>

namespace sr {

struct Lay {
    int nx, ny, nz;   // core nodes per direction
    int kh;           // used entries per half-row: (nz + 1) / 2 + 1
    int khs;          // half-row stride (elements)
};

__host__ __device__ inline long lay_size(const Lay &L) {
    return (long)(L.nx + 2) * (L.ny + 2) * 2 * L.khs;
}

__host__ __device__ inline int lay_color(int ip, int jp, int kp) { return (ip + jp + kp + 1) & 1; }

__host__ __device__ inline long lay_row(const Lay &L, int ip, int jp, int c) {
    return ((long)ip * (L.ny + 2) + jp) * 2L * L.khs + (long)c * L.khs;
}

__host__ __device__ inline long lay_idx(const Lay &L, int ip, int jp, int kp) {
    return lay_row(L, ip, jp, lay_color(ip, jp, kp)) + (kp >> 1);
}

// ghost rules of one direction (applied in the order x, y, z like imposeBC, poisson_d3.cc:575-668)
enum GhostRule : int {
    G_PERIODIC = 0,   // u(-1) = u(M-1), u(M+1) = u(1)          (updatePads self exchange)
    G_MIRROR = 1,     // u(-1) = u(1),   u(M+1) = u(M-1)          (Neumann)
    G_ANTI = 2,       // u(-1) = -u(1),  u(M+1) = -u(M-1)         (TEST_POISSON, zeroBC)
    G_WALL = 3,       // u(-1), u(M+1) = wall plane value         (TEST_POISSON, !zeroBC)
    G_ZERO = 4,       // ghost stays 0                           (residual pads, reference)
    G_KEEP = 5        // not touched                             (non-periodic updatePads)
};

struct Bc {
    int rule[3];
    const double *wall[6];   // x lo/hi, y lo/hi, z lo/hi (level-0 padded planes, device) or null
};

// source padded index of ghost padded index g (0 or n+1) of a direction with n nodes
__host__ __device__ inline int ghost_src(int rule, int g, int n) {
    if (rule == G_PERIODIC) return g == 0 ? n - 1 : 2;
    return g == 0 ? 2 : n - 1;   // mirror / anti
}

// Levels first..last (at most 8) of the coarse end of a V-cycle run by one workgroup in LDS.
// Offsets are in elements of the level type.
struct CoarseDesc {
    int first, last;
    Lay L[8];
    int off_u[8], off_f[8], off_c[8][3];
    const void *cg[8][3];
    int pre, post;
    Bc bc;       // smoothing ghosts of the coarse levels (zeroBC = 1)
    Bc tbc;      // residual pads read by the restriction
    double tol;
    int max_count;
    int project;
    const double *wc[3];
    double wcsum;
    const void *pinv;   // dense coarsest inverse, transposed (T), or null: Gauss-Seidel to tol
    int ncd;            // coarsest core nodes (rows of pinv)
    int off_s;          // scratch of the tap-parallel restriction, or -1 (none)
    int tap_max;        // coarse levels of at most this many core nodes use it
};

// Buffers of a two-stage final reduction (k_gcr_coef_mb): FINAL_STAGE doubles
// of per-workgroup sums and a device-scope counter that is 0 between launches.  nullptr in a
// launcher selects the single-workgroup final.
struct FinalBufs {
    double *stage;
    unsigned *counter;
};
constexpr int FINAL_STAGE = 16 * 8;

}  // namespace sr
