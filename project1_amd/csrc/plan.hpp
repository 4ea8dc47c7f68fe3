// Setup data of the wave-function-product kernels (the reference keeps the
// equivalent in module global / m_fft: gaunt1, integral, nfft, imati, gpt,
// gptm, cstep -- wavefproducts.f:157-222, fft.f:17-28).
#pragma once
#include <vector>

#include "common.hpp"

struct PlanMtType {
  int neq = 0, maxlm = 0, maxllm = 0;
  int llm0 = 0;  // first MT column of this type
  int ic0 = 0;   // first centre of this type
  double2 *T = nullptr;  // coupling tensor, (maxllm*maxlm) x maxlm complex
};

struct spexhip_plan {
  int kind = 0;  // 1 = muffin-tin, 2 = interstitial plane waves
  std::vector<void *> owned;  // device allocations released with the plan
  // muffin tin
  int maxlmindx = 0, nbasp = 0, ncent = 0;
  std::vector<PlanMtType> mt;
  std::vector<double> cent;
  // plane waves
  int nfft[3] = {0, 0, 0};
  int imati[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
  double vol = 1.0;
  int nkpt = 0, maxgpt = 0, ngptm = 0;
  std::vector<int> ngpt;          // per k
  std::vector<int> gptm;          // ngptm x 3 (mixed-basis IPWs of q)
  int *gidx = nullptr;            // device [nkpt][maxgpt] box index of each G
  double2 *step = nullptr;        // device theta(r) (complex) or null
  // momentum matrix (kind 3): per atom type the three Cartesian operators
  // B_j (nn x nn complex, column-major, device) that map the muffin-tin
  // coefficients of band 1 to the left vectors; the plane-wave basis of
  // every k (device G vectors) and the step function's coefficients
  struct MomType {
    int neq = 0, nn = 0;
    double2 *B = nullptr;  // 3 consecutive nn x nn matrices
  };
  std::vector<MomType> mom;
  int mom_ncent = 0;
  int *mom_gpt = nullptr;         // device [nkpt][maxgpt][3]
  int *mom_ngpt = nullptr;        // device [nkpt]
  double *mom_kpt = nullptr;      // device [nkpt][3], in the reciprocal basis
  double2 *mom_cstep = nullptr;   // device cstep, x fastest
  int cext[3] = {0, 0, 0};
  double rlat[9] = {0};           // rlat[i + 3*j]
};

// Box of Fourier components handled by the plane-wave kernels: each axis
// holds frequencies -8 .. 7 at index g + 8.
constexpr int PW_BOX = 16, PW_OFF = 8;
