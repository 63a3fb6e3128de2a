// Ghost and wall operations of one field as ordered waves (ts_kernels.h GhostWave), shared by the
// time step (ts_api.hip) and the field-operator entry points (field_api.hip).
#pragma once

#include "ts_kernels.h"

namespace sr {

inline void add_op(GhostWave &w, const GhostOp &op) {
    if (op.count > 0) w.ops[w.nops++] = op;
}

// mpidata::syncData (mpidata.cc:278-306) when the neighbour along periodic x / y is the rank itself:
// the faces over the core of the other two directions, then the four x-y edges over the z core.
// The edge exchange runs whatever the periodicity (edge ranks are never cleared, hydro.cc:174-212).
inline GhostWave sync_wave(const FShape &f, int xPer, int yPer) {
    GhostWave w{};
    int nx = f.n[0], ny = f.n[1], nz = f.n[2], sx = f.s[0], sy = f.s[1];
    auto plane = [&](int dim, int dst, int src, int alo, int ahi, int blo, int bhi) {
        GhostOp o{};
        o.kind = 0;
        o.dim = dim;
        o.dst[0] = dst;
        o.src[0] = src;
        o.lo[0] = alo; o.hi[0] = ahi;
        o.lo[1] = blo; o.hi[1] = bhi;
        o.a = 1.0; o.b = 0.0;
        o.count = (long)(ahi - alo + 1) * (bhi - blo + 1);
        add_op(w, o);
    };
    auto line = [&](int dx, int dy, int sxv, int syv) {
        GhostOp o{};
        o.kind = 1;
        o.dim = 2;
        o.dst[0] = dx; o.dst[1] = dy;
        o.src[0] = sxv; o.src[1] = syv;
        o.lo[0] = 0; o.hi[0] = nz - 1;
        o.a = 1.0; o.b = 0.0;
        o.count = nz;
        add_op(w, o);
    };
    if (xPer) {
        plane(0, -1, nx - 1 - sx, 0, ny - 1, 0, nz - 1);
        plane(0, nx, sx, 0, ny - 1, 0, nz - 1);
    }
    if (yPer) {
        plane(1, -1, ny - 1 - sy, 0, nx - 1, 0, nz - 1);
        plane(1, ny, sy, 0, nx - 1, 0, nz - 1);
    }
    line(-1, -1, nx - 1 - sx, ny - 1 - sy);
    line(-1, ny, nx - 1 - sx, sy);
    line(nx, -1, sx, ny - 1 - sy);
    line(nx, ny, sx, sy);
    return w;
}

// one wall of a field (lib/boundary/*.cc): the wall slice is the layer just outside the bulk
// (field.cc:180-290).  kind 0 periodicCC, 1 periodicFC, 2 dirichletCC, 3 dirichletFC,
// 4 neumannCC, 5 neumannFC.
inline GhostOp wall_op(const FShape &f, int per, int wall, int kind, double value) {
    int d = wall / 2, hi = wall & 1;
    int n = f.n[d];
    int blo = 0, bhi = n - 1;
    if (f.s[d] && !per) {
        blo = 1;
        bhi = n - 2;
    }
    int w = hi ? bhi + 1 : blo - 1;
    int wo = hi ? blo - 1 : bhi + 1;
    int src;
    switch (kind) {
        case 0: src = hi ? wo + 2 : wo - 2; break;
        case 1: src = hi ? wo + 1 : wo - 1; break;
        case 3: case 4: case 5: src = hi ? w - 1 : w + 1; break;
        default: src = w; break;
    }
    GhostOp o{};
    o.kind = 0;
    o.dim = d;
    o.dst[0] = w;
    o.src[0] = src;
    int a = d == 0 ? 1 : 0, b = d == 2 ? 1 : 2;
    o.lo[0] = -1; o.hi[0] = f.n[a];
    o.lo[1] = -1; o.hi[1] = f.n[b];
    o.count = (long)(f.n[a] + 2) * (f.n[b] + 2);
    switch (kind) {
        case 2: o.a = 0.0; o.b = value; break;
        case 3: o.a = -1.0; o.b = 2.0 * value; break;     // 2 v - F(wall +- 1)
        default: o.a = 1.0; o.b = 0.0; break;
    }
    return o;
}

}  // namespace sr
