// Kleinman-Bylander projectors of one k-point built on the device (FP64), the
// projector construction of Src/omp/proj_nl_kb_c16.f90:23-275:
//
//   anlga(i, ind) = exp(-2 pi i g_i . rat) rylm_lm(q_i) i^l vkb_l(|q_i|) / sqrt(vcell)
//
// with q_i = bvec (k + g_i) Cartesian (:194-203), rylm the real harmonics
// times |q|^l sqrt(4 pi / (2l+1)) of Src/numeric/ylm_rc.f90:82-107, vkb_l the
// quadratic interpolation of the type's table at |q| / delqnl, zero from
// nqnl - 1 on (:227-240), and columns ordered type -> l -> m -> atom (:169-186).
//
// One thread per (plane wave i, atom): the phase is computed once per atom
// (sincospi of -2 g.rat) and reused by the 2l+1 columns of every l of the
// atom's type; for a fixed column a wave stores 64 consecutive rows (1 KB),
// so the 16 bytes per element written are the whole HBM traffic besides the
// 24-byte g_i read once per atom.
#pragma once

#include <hip/hip_runtime.h>

namespace cpwgpu {

struct KbType {
    int nq;          // nqnl: table points; interpolation needs nint(|q|/delq) < nq - 1
    int lmask;       // bit l set when nkb(l, 0, type) /= 0
    int natom;       // atoms of the type
    int base[4];     // first column of (type, l): column = base[l] + (m + l) * natom + atom
    double inv_delq; // 1 / delqnl
    long voff;       // offset of the type's l = 0 line in the packed table (line l at voff + l * vstride)
};

struct KbAtom {
    double rat[3];   // lattice coordinates
    int type, atom;
};

struct KbCell {
    double b[9];     // bvec column-major: b[r + 3 c] = bvec(r+1, c+1)
    double rk[3];    // k-point, reciprocal lattice coordinates
    double fac;      // 1 / sqrt(vcell)
};

// rylm_lm(x, y, z) of Src/numeric/ylm_rc.f90:82-107, l <= 3, m = -l..l
__device__ __forceinline__ double kb_rylm(int l, int m, double x, double y, double z) {
    const double s3 = 1.7320508075688772935;   // sqrt(3)
    const double s34 = 0.86602540378443864676; // sqrt(3/4)
    const double s58 = 0.79056941504209483300; // sqrt(5/8)
    const double s15 = 3.8729833462074168852;  // sqrt(15)
    const double s38 = 0.61237243569579452455; // sqrt(3/8)
    const double s154 = 1.9364916731037084426; // sqrt(15/4)
    switch (l * 8 + m + 3) {
    case 0 * 8 + 3: return 1.0;
    case 1 * 8 + 2: return y;
    case 1 * 8 + 3: return z;
    case 1 * 8 + 4: return x;
    case 2 * 8 + 1: return s3 * x * y;
    case 2 * 8 + 2: return s3 * z * y;
    case 2 * 8 + 3: return 0.5 * (2.0 * z * z - x * x - y * y);
    case 2 * 8 + 4: return s3 * z * x;
    case 2 * 8 + 5: return s34 * (x * x - y * y);
    case 3 * 8 + 0: return s58 * (3.0 * x * x - y * y) * y;
    case 3 * 8 + 1: return s15 * z * x * y;
    case 3 * 8 + 2: return s38 * (4.0 * z * z - x * x - y * y) * y;
    case 3 * 8 + 3: return 0.5 * (2.0 * z * z - 3.0 * x * x - 3.0 * y * y) * z;
    case 3 * 8 + 4: return s38 * (4.0 * z * z - x * x - y * y) * x;
    case 3 * 8 + 5: return s154 * (x * x - y * y) * z;
    case 3 * 8 + 6: return -s58 * (3.0 * y * y - x * x) * x;
    default: return 0.0;
    }
}

constexpr int KB_THREADS = 256;

// grid (ceil(mtxd / KB_THREADS), number of atoms); gw[j * mtxd + i] = g_j of
// plane wave i; vkb packed per (type, l) line, line[j + 2] = vkb(j).  Every
// table read is at line position nint(|q|/delq) + {-1, 0, 1} + 2 with
// nint(|q|/delq) in 0..nq - 2, so in 1..nq + 1: inside the line (vstride
// values) when nq <= vstride - 2 (checked on the host).
__global__ __launch_bounds__(KB_THREADS) void k_proj_kb(const double *__restrict__ gw, int mtxd, KbCell cell,
                                                        const KbAtom *__restrict__ atoms,
                                                        const KbType *__restrict__ types,
                                                        const double *__restrict__ vkb, long vstride,
                                                        double2 *__restrict__ anl, long lda) {
    const int i = blockIdx.x * KB_THREADS + threadIdx.x;
    if (i >= mtxd) return;
    const KbAtom at = atoms[blockIdx.y];
    const KbType ty = types[at.type];
    const double g1 = gw[i], g2 = gw[(long)mtxd + i], g3 = gw[2L * mtxd + i];
    const double k1 = cell.rk[0] + g1, k2 = cell.rk[1] + g2, k3 = cell.rk[2] + g3;
    const double x = cell.b[0] * k1 + cell.b[3] * k2 + cell.b[6] * k3;
    const double y = cell.b[1] * k1 + cell.b[4] * k2 + cell.b[7] * k3;
    const double z = cell.b[2] * k1 + cell.b[5] * k2 + cell.b[8] * k3;
    const double qi = sqrt(x * x + y * y + z * z);
    // structure phase exp(-2 pi i g.rat), argument reduced in units of pi
    double sn, cs;
    sincospi(-2.0 * (g1 * at.rat[0] + g2 * at.rat[1] + g3 * at.rat[2]), &sn, &cs);
    // interpolation point; a NaN or out-of-table index gives zero, as past nqnl - 1
    const double xni = qi * ty.inv_delq;
    const double nid = floor(xni + 0.5);
    const bool inside = nid >= 0.0 && nid < (double)(ty.nq - 1);
    const int ni = inside ? (int)nid : 0;
    const double xi = xni - (double)ni;
    for (int l = 0; l <= 3; l++) {
        if (!(ty.lmask >> l & 1)) continue;
        double vq = 0.0;
        if (inside) {
            const double *line = vkb + ty.voff + (long)l * vstride + 2;
            vq = line[ni] * (1.0 + xi) * (1.0 - xi) + 0.5 * (line[ni + 1] * (1.0 + xi) - line[ni - 1] * (1.0 - xi)) * xi;
            vq *= cell.fac;
        }
        // st * vq * i^l
        double re = cs * vq, im = sn * vq;
        for (int r = 0; r < l; r++) { const double t = re; re = -im; im = t; }
        double2 *col = anl + (long)(ty.base[l] + at.atom) * lda + i;
        for (int m = -l; m <= l; m++) {
            const double a = kb_rylm(l, m, x, y, z);
            col[(long)(m + l) * ty.natom * lda] = make_double2(re * a, im * a);
        }
    }
}

}  // namespace cpwgpu
