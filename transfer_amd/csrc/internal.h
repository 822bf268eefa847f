// internal.h -- host-side structures shared by the transfer_amd sources.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "transfer.h"

namespace tr {

// Sticky per-thread error text behind tr_last_error().
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

// SplitMix64 counter-based generator: u(seed, stream, i) in [0, 1).
uint64_t splitmix64(uint64_t x);
double rng_uniform(uint64_t seed, uint64_t stream, uint64_t i);
double rng_normal(uint64_t seed, uint64_t stream, uint64_t i);

// ---------------------------------------------------------------------------
// AngularData (AngularData.h:8-18) plus the scattering extension.
struct Angular {
    int maxk = 0;
    uint32_t slm = 0;
    std::vector<double> omega;       // [3][slm][slm] (reference storage)
    std::vector<uint32_t> omega_pos; // [3][slm][slm]
    std::vector<double> Ox, Oy, Oz;  // [slm][slm]
    std::vector<double> massfac;     // [slm]
    std::vector<int> l_of, m_of;     // harmonic (l, m) of each lm
};
void build_angular(int maxk, double albedo, double g, int hg_lmax, Angular &a);

// Real spherical harmonics of Spherical.cpp's convention for all l <= lmax,
// out[l*l + l + m].
void sh_eval_all(int lmax, double x, double y, double z, double *out);
// Gauss-Legendre nodes/weights on [-1, 1].
void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w);

// ---------------------------------------------------------------------------
// MeshData (MeshData.h:13-32) built from raw geometry.
struct Mesh {
    uint32_t nP = 0, nT = 0, nF = 0;
    std::vector<double> coords;      // [nP][3]
    std::vector<uint32_t> tets;      // [nT][4]
    std::vector<double> kappa, Ip;   // [nT]
    std::vector<double> volume;      // [nT]
    std::vector<uint32_t> tetstart, tetidx, tetpos;
    std::vector<uint32_t> facestart, faceidx, facepos;
    std::vector<tr_tetrahedron> mesh;
    std::vector<tr_face> bnd;
    std::vector<uint8_t> face_kind;  // [nF]
    std::vector<uint32_t> face_tet;  // [nF] owning tet
    std::vector<uint32_t> region;    // [nT] material number of a mesh file, empty for generated meshes
};
int build_mesh(uint32_t nP, const double *coords, uint32_t nT,
               const uint32_t *tets, const double *kappa, const double *Ip,
               Mesh &m);
void refresh_fields(Mesh &m);
int generate_box(const tr_mesh_spec &spec, Mesh &m);

// ---------------------------------------------------------------------------
// Config (Config.h:4-48)
struct Config {
    int nreg = -1, maxk = 1, dev = 0, dump = 0, prec = 0, soltype = 0;
    std::vector<double> kappa, Ip;
    std::string meshfn, outfn = "solution.vtk";
};
int load_config(const char *path, Config &c);

}  // namespace tr
