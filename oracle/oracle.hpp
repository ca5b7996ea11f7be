// CPU oracle of the particle hot path: TEST INFRASTRUCTURE ONLY.
//
// A restatement of the reference's kernels, written from the reference
// source read as text (the reference itself is CUDA 8 / sm_20 /
// Windows-only and cannot be built here).  Every function follows one
// reference kernel, in its loop order (27 cells in neighborid order, members
// ascending) and branch structure, and cites the lines it restates.  It is
// templated on the arithmetic type: float is the faithful FP32 restatement,
// double is the FP64 twin used to report every result's error.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
// load liboracle.so, and only as the checker or the CPU baseline; the
// product (libsphx.so) never links or calls it.
#pragma once

#include <cstdint>

#include "../include/sphx.h"

extern "C" {

// Pointers to caller-owned arrays (numpy in the tests).  Vectors are
// [n][3], matrices [n][9] row-major; T is float (orc_f32_*) or double
// (orc_f64_*); integer arrays are shared by both.
typedef struct OrcState {
  int n;
  int gridtotal;
  void* pos;        /* T[n][3] */
  int* type;        /* int[n]  */
  void* vel;        /* T[n][3] */
  void* veleval;    /* T[n][3] */
  void* accel;      /* T[n][3] */
  void* pressure;   /* T[n]    */
  void* dens;
  void* restdens;
  void* mass;
  void* visc;
  int* bornid;
  /* grid */
  uint32_t* mgcell;
  uint32_t* midsort;
  uint32_t* mgndx;
  int* gridoff;     /* int[gridtotal] */
  int* gridcnt;
  int* MFidTable;
  /* IISPH */
  void* dii;        /* T[n][3] */
  void* dijpj;      /* T[n][3] */
  void* aii;        /* T[n] */
  void* rho_adv;
  void* press_l;
  void* press_l1;
  void* residue;
} OrcState;

}  // extern "C"
