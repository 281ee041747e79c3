// Internal types shared by the host setup (ajet_setup.cpp) and the gfx950 kernels
// (ajet_kernels.hip).  Not part of the C ABI.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/ajet_capi.h"

namespace ajet {

// Tile of the fused stage kernel: TZ x TR output points, halo 2 on every side.
#ifndef AJET_TZ
#define AJET_TZ 32
#endif
#ifndef AJET_TR
#define AJET_TR 16
#endif
#ifndef AJET_THREADS
#define AJET_THREADS 256
#endif
constexpr int TZ = AJET_TZ;
constexpr int TR = AJET_TR;
constexpr int HALO = 2;
constexpr int WZ = TZ + 2 * HALO;
constexpr int WR = TR + 2 * HALO;
constexpr int STAGE_THREADS = AJET_THREADS;

constexpr int LAD_LINES = 64;  // r-lines one wave of the fused z/r filter walks
constexpr int LAD_REACH = 7;   // points beyond a source tile at which it reads |S| and dil

constexpr int MAX_PATCH = 64;
constexpr int MAX_SEQ = 64;       // distinct ordered lists of boundary patches met at one point
constexpr int SEQ_LEN = 6;

// Per-point metadata word:
//   bits 0-3  SBP row of the z derivative (4 = interior, 0-3 / 5-8 closure rows, 15 = blanked)
//   bits 4-7  SBP row of the r derivative
//   bits 8-15 index of the ordered boundary-patch list applied at the point (0 = none)
constexpr uint32_t ROW_INTERIOR = 4;
constexpr uint32_t ROW_BLANK = 15;

struct PatchDev {
  int32_t bc, dir;
  int32_t i_start, i_end, j_start, j_end;  // 1-based, inclusive
};

// Scalars the step keeps on the device so that no step needs the host.
struct StepScalars {
  double dt;              // delt of the current step (ComputeTimeStep result)
  double time;            // accumulated time (ModRungeKutta.f90:152-155)
  double T_stale;         // the module scalar T of ModInputOutput.f90:404-405
  unsigned long long dmax_bits;  // max over points of the CFL denominator (bit pattern)
  long long nlog;         // entries written to the time-step log
  int32_t nonfinite;      // a non-finite denominator was met
  int32_t pad;
  double tpar[2];         // time at the start of the step, by step parity: the step reads one
                          // slot and its last stage writes the other, so no launch reads a
                          // value it also writes
};

struct Geometry {
  int nz, nr, pitch;      // pitch: row stride (elements) of every field on the device
  long long fstride;      // field stride = pitch * nr
  int ntz, ntr;           // tiles along z and r
  int any_blank;
  int iT, jT;             // 0-based point whose T becomes the stale scalar after a step (-1: none)
  int nseq;
  int jtop;               // first row whose r reaches the top sponge (nr: none)
};

// Host-side setup products (pure CPU, shared with tests through the C ABI).
struct Setup {
  std::vector<ajet_patch> patches;
  std::vector<uint32_t> meta;           // (nz, nr)
  std::vector<double> zc, rc;           // 1-D coordinates
  std::vector<double> dxidz, ddxidxidz; // (nz)
  std::vector<double> detadr, ddetadetadr;  // (nr)
  std::vector<uint8_t> seq;             // MAX_SEQ x SEQ_LEN patch indices (1-based, 0 = end)
  int nseq = 1;
  int iT = -1, jT = -1;
  int any_blank = 0;
  std::vector<int> zcuts, rcuts;        // tile boundaries along z and r (0, ..., n)
};

void sbp_tables(double sbp1[9][6], double sbp2[9][6], double tg[9][11]);
int build_patches(int nz, int nr, const ajet_bc* lines, int nlines, std::vector<ajet_patch>& out);
int build_setup(const ajet_params& p, const double* z, const double* r, const int32_t* iblank,
                const ajet_bc* lines, int nlines, Setup& s, std::string& err);
int build_tiling(Setup& s, int nz, int nr, std::string& err);
void classify_tiles(const Setup& s, int nz, int nr, std::vector<int>& interior, std::vector<int>& boundary);
void classify_lad_tiles(const Setup& s, int nz, int nr, std::vector<int>& interior, std::vector<int>& boundary);
void classify_lad_source(const Setup& s, int nz, int nr, std::vector<int>& fused, std::vector<int>& other,
                         std::vector<int>& str_interior, std::vector<int>& str_boundary);
void mixed_tile_list(const std::vector<int>& interior, const std::vector<int>& boundary, std::vector<int>& out);
void build_lad_line_groups(const Setup& s, int nz, int nr, int lines, std::vector<int>& groups);
void lad_tables(const Setup& s, int nz, int nr, std::vector<double>& dz, std::vector<double>& dr,
                std::vector<double>& ramp);

// learned LAD coefficient: BF16 rounding (nearest even) and the device layout of a model
uint16_t bf16_bits(float f);
void mlp_pack(const ajet_mlp& m, std::vector<uint16_t>& W, std::vector<float>& bias);

}  // namespace ajet
