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

template <class T, int ALG, int SAVE, bool KAHAN, int WAVES = 0>
const void* adaptive_ptr() {
  return reinterpret_cast<const void*>(&sdeh_adaptive_kernel<T, R, ALG, SAVE, KAHAN, WAVES>);
}

// Register-cap variants (SDEH_WAVE_VARIANTS, built for the Lorenz right-hand side):
// the adaptive endpoint kernels of ATsit5 and AVern7 at 4, 6 and 8 waves per SIMD,
// for the autotuner; waves = 0 is the default kernel.

template <class T, int ALG, bool KAHAN>
const void* waves_variant(int waves) {
#if defined(SDEH_WAVE_VARIANTS)
  if constexpr (ALG == sdeh::ALG_ATSIT5 || ALG == sdeh::ALG_AVERN7) {
    switch (waves) {
      case 4: return adaptive_ptr<T, ALG, sdeh::SAVE_END, KAHAN, 4>();
      case 6: return adaptive_ptr<T, ALG, sdeh::SAVE_END, KAHAN, 6>();
      case 8: return adaptive_ptr<T, ALG, sdeh::SAVE_END, KAHAN, 8>();
    }
  }
#endif
  (void)waves;
  return nullptr;
}

// Two trajectories per lane (SDEH_WAVE_VARIANTS): ATsit5 (binary64, binary32 plain
// and compensated) and AVern7 (binary64), endpoint output.
template <class T, int ALG, bool KAHAN>
const void* multi_variant(int tpl) {
#if defined(SDEH_WAVE_VARIANTS)
  if constexpr ((ALG == sdeh::ALG_ATSIT5) || (ALG == sdeh::ALG_AVERN7 && sizeof(T) == 8 && !KAHAN)) {
    if (tpl == 2)
      return reinterpret_cast<const void*>(&sdeh_adaptive_multi_kernel<T, R, ALG, KAHAN, 2>);
  }
#endif
  (void)tpl;
  return nullptr;
}

// Which variants this translation unit instantiates.  The 16-species GLV
// right-hand side is only built for the variants its benchmark and tests use
// (each of its kernels is large); SDEH_LIMITED selects that subset.
template <class T, int ALG, int SAVE>
constexpr bool allowed() {
#if defined(SDEH_WAVE_ONLY)
  // The matrix-core layout: binary64 AVern9 and ATsit5, binary32 ATsit5 (plain and
  // compensated), endpoint and saveat output.
  return (sizeof(T) == 8 && (ALG == sdeh::ALG_AVERN9 || ALG == sdeh::ALG_ATSIT5) &&
          SAVE != sdeh::SAVE_STEPS) ||
         (sizeof(T) == 4 && ALG == sdeh::ALG_ATSIT5 && SAVE != sdeh::SAVE_STEPS);
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
const void* by_save(int save, int kahan, int waves, int tpl) {
  constexpr bool fixed = (ALG == sdeh::ALG_TSIT5 || ALG == sdeh::ALG_VERN7 ||
                          ALG == sdeh::ALG_VERN9 || ALG == sdeh::ALG_RK4 ||
                          ALG == sdeh::ALG_EULER);
  // GPUSimpleRK4 and GPUSimpleEuler have no interpolant: no saveat kernel.
  constexpr bool dense = !(ALG == sdeh::ALG_RK4 || ALG == sdeh::ALG_EULER);
  if constexpr (fixed) {
    if (kahan || waves || tpl > 1) return nullptr;
    switch (save) {
      case sdeh::SAVE_END: return fixed_or_null<T, ALG, sdeh::SAVE_END>();
      case sdeh::SAVE_AT:
        if constexpr (dense) return fixed_or_null<T, ALG, sdeh::SAVE_AT>();
        return nullptr;
      case sdeh::SAVE_STEPS: return fixed_or_null<T, ALG, sdeh::SAVE_STEPS>();
    }
    return nullptr;
  } else {
    if (tpl > 1) {
      if (save != sdeh::SAVE_END || waves) return nullptr;
      if (kahan) {
        if constexpr (sizeof(T) == 4) return multi_variant<T, ALG, true>(tpl);
        return nullptr;
      }
      return multi_variant<T, ALG, false>(tpl);
    }
    if (waves) {
      if (save != sdeh::SAVE_END) return nullptr;
      if (kahan) {
        if constexpr (sizeof(T) == 4) return waves_variant<T, ALG, true>(waves);
        return nullptr;
      }
      return waves_variant<T, ALG, false>(waves);
    }
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
const void* by_alg(int alg, int save, int kahan, int waves, int tpl) {
  switch (alg) {
    case sdeh::ALG_TSIT5: return by_save<T, sdeh::ALG_TSIT5>(save, kahan, waves, tpl);
    case sdeh::ALG_ATSIT5: return by_save<T, sdeh::ALG_ATSIT5>(save, kahan, waves, tpl);
    case sdeh::ALG_VERN7: return by_save<T, sdeh::ALG_VERN7>(save, kahan, waves, tpl);
    case sdeh::ALG_AVERN7: return by_save<T, sdeh::ALG_AVERN7>(save, kahan, waves, tpl);
    case sdeh::ALG_VERN9: return by_save<T, sdeh::ALG_VERN9>(save, kahan, waves, tpl);
    case sdeh::ALG_AVERN9: return by_save<T, sdeh::ALG_AVERN9>(save, kahan, waves, tpl);
    case sdeh::ALG_RK4: return by_save<T, sdeh::ALG_RK4>(save, kahan, waves, tpl);
    case sdeh::ALG_EULER: return by_save<T, sdeh::ALG_EULER>(save, kahan, waves, tpl);
  }
  return nullptr;
}

}  // namespace

// dtype: 0 binary64, 1 binary32; waves: 0 for the default kernel, else a
// register-cap variant; tpl: trajectories per lane (1, or 2 for the multi-slot
// variant).  Returns the kernel's host stub, or null where not built.
extern "C" __attribute__((visibility("hidden"))) const void* SDEH_AOT_FN(int dtype, int alg,
                                                                         int save, int kahan,
                                                                         int waves, int tpl) {
  if (dtype == 0) return by_alg<double>(alg, save, kahan, waves, tpl);
  if (dtype == 1) return by_alg<float>(alg, save, kahan, waves, tpl);
  return nullptr;
}
