// Ahead-of-time integrator kernels for one built-in right-hand side.
// Compiled once per RHS with -DSDEH_RHS=<struct in sdeh_rhs.h> and
// -DSDEH_AOT_FN=<lookup function name>; the runtime looks kernels up through
// that function.
#include <hip/hip_runtime.h>

#include "sdeh_integrators.h"
#include "sdeh_rhs.h"

#ifndef SDEH_RHS
#error "define SDEH_RHS"
#endif

namespace {

using R = sdeh_rhs::SDEH_RHS;

template <class T, int ALG, int SAVE>
const void* fixed_ptr() {
  return reinterpret_cast<const void*>(&sdeh_fixed_kernel<T, R, ALG, SAVE>);
}

template <class T, int ALG, int SAVE, bool KAHAN>
const void* adaptive_ptr() {
  return reinterpret_cast<const void*>(&sdeh_adaptive_kernel<T, R, ALG, SAVE, KAHAN>);
}

// Which variants this translation unit instantiates.  The 16-species GLV
// right-hand side is only built for the variants its benchmark and tests use
// (each of its kernels is large); SDEH_LIMITED selects that subset.
template <class T, int ALG, int SAVE>
constexpr bool allowed() {
#if defined(SDEH_WAVE_ONLY)
  // The matrix-core layout: binary64 adaptive solvers, endpoint and saveat output.
  return sizeof(T) == 8 && (ALG == sdeh::ALG_AVERN9 || ALG == sdeh::ALG_ATSIT5) &&
         SAVE != sdeh::SAVE_STEPS;
#elif defined(SDEH_LIMITED)
  return (ALG == sdeh::ALG_AVERN9 && sizeof(T) == 8 && SAVE != sdeh::SAVE_STEPS) ||
         (ALG == sdeh::ALG_ATSIT5 && SAVE != sdeh::SAVE_STEPS) ||
         (ALG == sdeh::ALG_TSIT5 && SAVE == sdeh::SAVE_END);
#else
  return true;
#endif
}

template <class T, int ALG, int SAVE>
const void* fixed_or_null() {
  if constexpr (allowed<T, ALG, SAVE>()) return fixed_ptr<T, ALG, SAVE>();
  return nullptr;
}

template <class T, int ALG, int SAVE, bool KAHAN>
const void* adaptive_or_null() {
  if constexpr (allowed<T, ALG, SAVE>()) return adaptive_ptr<T, ALG, SAVE, KAHAN>();
  return nullptr;
}

template <class T, int ALG>
const void* by_save(int save, int kahan) {
  constexpr bool fixed = (ALG == sdeh::ALG_TSIT5 || ALG == sdeh::ALG_VERN7 ||
                          ALG == sdeh::ALG_VERN9 || ALG == sdeh::ALG_RK4 ||
                          ALG == sdeh::ALG_EULER);
  // GPUSimpleRK4 and GPUSimpleEuler have no interpolant: no saveat kernel.
  constexpr bool dense = !(ALG == sdeh::ALG_RK4 || ALG == sdeh::ALG_EULER);
  if constexpr (fixed) {
    if (kahan) return nullptr;
    switch (save) {
      case sdeh::SAVE_END: return fixed_or_null<T, ALG, sdeh::SAVE_END>();
      case sdeh::SAVE_AT:
        if constexpr (dense) return fixed_or_null<T, ALG, sdeh::SAVE_AT>();
        return nullptr;
      case sdeh::SAVE_STEPS: return fixed_or_null<T, ALG, sdeh::SAVE_STEPS>();
    }
    return nullptr;
  } else {
    if (kahan) {
      // The compensated update is a binary32 precision mode.
      if constexpr (sizeof(T) == 4) {
        if (save == sdeh::SAVE_END) return adaptive_or_null<T, ALG, sdeh::SAVE_END, true>();
      }
      return nullptr;
    }
    switch (save) {
      case sdeh::SAVE_END: return adaptive_or_null<T, ALG, sdeh::SAVE_END, false>();
      case sdeh::SAVE_AT: return adaptive_or_null<T, ALG, sdeh::SAVE_AT, false>();
      case sdeh::SAVE_STEPS: return adaptive_or_null<T, ALG, sdeh::SAVE_STEPS, false>();
    }
    return nullptr;
  }
}

template <class T>
const void* by_alg(int alg, int save, int kahan) {
  switch (alg) {
    case sdeh::ALG_TSIT5: return by_save<T, sdeh::ALG_TSIT5>(save, kahan);
    case sdeh::ALG_ATSIT5: return by_save<T, sdeh::ALG_ATSIT5>(save, kahan);
    case sdeh::ALG_VERN7: return by_save<T, sdeh::ALG_VERN7>(save, kahan);
    case sdeh::ALG_AVERN7: return by_save<T, sdeh::ALG_AVERN7>(save, kahan);
    case sdeh::ALG_VERN9: return by_save<T, sdeh::ALG_VERN9>(save, kahan);
    case sdeh::ALG_AVERN9: return by_save<T, sdeh::ALG_AVERN9>(save, kahan);
    case sdeh::ALG_RK4: return by_save<T, sdeh::ALG_RK4>(save, kahan);
    case sdeh::ALG_EULER: return by_save<T, sdeh::ALG_EULER>(save, kahan);
  }
  return nullptr;
}

}  // namespace

// dtype: 0 binary64, 1 binary32.  Returns the kernel's host stub or null.
extern "C" __attribute__((visibility("hidden"))) const void* SDEH_AOT_FN(int dtype, int alg,
                                                                         int save, int kahan) {
  if (dtype == 0) return by_alg<double>(alg, save, kahan);
  if (dtype == 1) return by_alg<float>(alg, save, kahan);
  return nullptr;
}
