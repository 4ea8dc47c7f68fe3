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
  double *step = nullptr;         // device theta(r) or null
};

// Box of Fourier components handled by the plane-wave kernels: each axis
// holds frequencies -8 .. 7 at index g + 8.
constexpr int PW_BOX = 16, PW_OFF = 8;
