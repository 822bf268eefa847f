// stencil.h -- the device operator's data, assembled on the host in FP64.
//
// The reference applies A vertex by vertex, walking each vertex's incident
// tets and rebuilding every tet's gradient once per vertex
// (kernels.cu:200-287).  Here the spatial part is assembled once into
// per-edge couplings ("channels") and the angular part stays matrix-free:
//
//   (Af)_i = sum_k massfac * M_ik f_k
//          + sum_c Omega_c ( sum_k K^c_ik f_k )            c = 1..NC-1
//          + sum_n O_n ( sum_k B^n_ik f_k )                n = x, y, z
//
//   M_ik   = sum_{t ∋ i,k} kappa V / 20 (1 + delta_ik)                  (kernels.cu:249)
//   K^ab_ik= sum_{t ∋ i,k} s^a_{t,i} s^b_{t,k} / (9 kappa V)            (kernels.cu:251-260)
//   channels 1..3 = K^xx, K^yy, K^zz;  4..6 = K^ab + K^ba (xy, xz, yz);
//   7..9 = K^ab - K^ba (only when the angular matrices are not symmetric
//   in a <-> b, i.e. with anisotropic scattering)
//   B^n_ik = sum over vacuum faces F ∋ i,k with axis n of S_F/12 (1 + delta_ik)
//            (kernels.cu:414-435, axis rule :419-425)
//
// Vertices are renumbered (Morton order of coordinates, or reverse
// Cuthill-McKee without coordinates) so that the neighbour gathers of a
// wavefront hit nearby memory.  Neighbour lists are ELL (fixed width W,
// slot 0 = the vertex itself, unused slots point at the vertex with zero
// coefficients) stored slot-major so that a wave's vertices read adjacent
// words.
#pragma once

#include <cstdint>
#include <vector>

#include "transfer.h"

namespace tr {

struct HostOperator {
    uint32_t nP = 0, slm = 0, aslm = 0, NC = 7, W = 0;
    std::vector<uint32_t> perm;   // perm[new] = caller vertex id
    std::vector<uint32_t> iperm;  // iperm[caller id] = new
    std::vector<uint32_t> nbr;    // [W][nP]
    std::vector<double> coef;     // [W][NC][nP]
    // vacuum-boundary stencil over the boundary vertices only
    uint32_t nB = 0, WB = 0;
    std::vector<uint32_t> bvert;  // [nB]
    std::vector<int32_t> bof;     // [nP] boundary row of each vertex, -1 inside
    std::vector<uint32_t> bnbr;   // [WB][nB]
    std::vector<double> bcoef;    // [WB][3][nB]
    // angular tables: for output lm, the padded list of (channel, lm')
    // couplings; idx = (c-1)*slm + lm' (volume) or n*slm + lm' (boundary)
    uint32_t arow = 0, brow = 0;
    std::vector<uint16_t> aidx, bidx;  // [row][slm]
    std::vector<double> aval, bval;    // [row][slm]
    std::vector<double> massfac;       // [slm]
    std::vector<double> ddiag;         // [3][slm]  Omega^aa[lm][lm] (kernels.cu:348-353)
    std::vector<double> bdiag;         // [3][slm]  O_n[lm][lm]      (kernels.cu:490)
    std::vector<double> rhs;           // [nP] sum kappa V I_p / 4   (kernels.cu:157-163)
    // dense channel matrices Omega_c[out][in], c = 1..NC-1, and O_n[out][in]
    std::vector<double> omc;           // [NC-1][slm][slm]
    std::vector<double> onc;           // [3][slm][slm]
    // specular reflection on axis-aligned planes (extension): entries
    // v*slm + lm of the internal [v][slm] layout that vanish, i.e. the
    // harmonics odd under the mirror of a reflective plane through v
    std::vector<uint32_t> mask;
    // first column of the reference's compressed angular data (omega[i][lm][0],
    // omega_pos[i][lm][0], i = 0..2) for the moments T_ij (MeshData.cpp:96-116)
    std::vector<double> mom_val;    // [3][slm]
    std::vector<uint32_t> mom_pos;  // [3][slm]
    // batches: per-vertex factor of the lm = 0 mass term (1 - albedo of the
    // vertex's problem, isotropic scattering); empty outside batches
    std::vector<double> mf0;
    // inputs of the learned guess's per-vertex features, in the new
    // numbering (empty when the mesh has no coordinates, and on coarse
    // levels): vsum[0..2][nP] = sums over the incident tets of V, kappa V
    // and kappa V I_p; xyz[3][nP] = the vertex coordinates
    std::vector<double> vsum, xyz;
    // vacuum faces whose normal is not along an axis, with their exact
    // angular matrices (extension; only when general_normals is set before
    // assembly, else such faces take the reference's axis rule):
    //   r_i += O(n_F) sum_{k in F} (S_F / 12)(1 + delta_ik) f_k
    // gvert[ngv] the vertices with such faces, gstart[ngv+1] into
    // gface/gpos (face, the vertex's place in it), gfv[nGF][3] the faces'
    // vertices, garea[nGF] = S_F / 12.  <|Omega.n| Y Y'> = R O_z R^T, with
    // gR[nGF][rotation_size] the faces' rotations (normal_rotation), gOz
    // [slm][slm] = O_z and gOd[nGF][slm] the diagonals of R O_z R^T
    bool general_normals = false;
    std::vector<uint32_t> gvert, gstart, gface, gpos, gfv;
    std::vector<double> garea, gR, gOd, gOz;
};

// <|Omega.n| Y_out Y_in> for the unit normal n: [slm][slm] (symmetric),
// exact (a Gauss-Legendre x trapezoid rule on the hemisphere around n; the
// integrand is even, and a polynomial of degree 4K+1 there).
void abs_normal_matrix(int maxk, const double n[3], std::vector<double> &O);
// column 00 of abs_normal_matrix(n) (slm values): <|Omega.n| Y_lm>
void abs_normal_column(int maxk, const double n[3], double *col);

// The rotation of the even harmonics up to 2K that takes the z axis to n
// (the frame of abs_normal_matrix): R_l[a][b] = <Y_la(Q w) Y_lb(w)>, Q the
// frame, one (2l+1)^2 block per even l, packed (rotation_size values).
// abs_normal_matrix(n) = R O_z R^T with O_z = abs_normal_matrix(z axis).
uint32_t rotation_size(int maxk);
void normal_rotation(int maxk, const double n[3], std::vector<double> &R);

// The block-diagonal union of independent operators with identical angular
// data (the problems of a batch), vertices of problem p at
// voff[p] .. voff[p+1] in the union's internal numbering and caller ids
// shifted by the same offsets.  mf0 gets 1 - albedo[p] per vertex.
// Galerkin coarse operator under aggregation of runs of `size` consecutive
// vertices (per problem range voff, or all vertices when voff is empty):
// agg[I] .. agg[I+1] are aggregate I's fine vertices, cvoff the coarse
// problem ranges (coarsen.cpp).
int coarsen_operator(const HostOperator &f, const std::vector<uint32_t> &voff, uint32_t size, HostOperator &c,
                     std::vector<uint32_t> &agg, std::vector<uint32_t> &cvoff);

int concat_operators(const std::vector<HostOperator> &ops, const std::vector<double> &albedo, HostOperator &u,
                     std::vector<uint32_t> &voff);

// Bit a of mirror_odd(l, m) is set when the real harmonic Y_lm (cos(m phi)
// for m >= 0, sin(|m| phi) for m < 0, polar axis z, phi from x) changes sign
// under the mirror x_a -> -x_a.
unsigned mirror_odd(int l, int m);

int assemble_operator(const tr_mesh_view &mv, const tr_angular_view &av,
                      HostOperator &op);

}  // namespace tr
