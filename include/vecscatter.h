/**
 * @file vecscatter.h
 * @brief Element type and stencil names for the brick headers.
 *
 * In bricklib (include/vecscatter.h:12-14) `bElem` is the element type,
 * `double` unless defined before inclusion, and `brick(...)` / `tile(...)`
 * are pragmas that a Python code generator replaces with vector code.  Here
 * there is no code generator: a stencil file name selects one of the
 * library's hand-written kernels (bk_stencil_from_file) and a whole sweep
 * over a list of bricks runs on the GPU through brick_sweep() in
 * brick-hip.h.  The per-brick `brick(...)` statement therefore becomes one
 * call after the loop that used to contain it.
 */
#ifndef BK_VECSCATTER_H
#define BK_VECSCATTER_H

#include <cstring>

#include "bk_api.h"

#ifndef bElem
#define bElem double
#endif

/// The reference's vector-ISA selector; only the HIP path exists here.
#ifndef VSVEC
#define VSVEC "HIP"
#endif

/// Stencil kind of a reference stencil file name ("7pt.py", "cond.py",
/// "mpi13pt.py", "mpi25pt.py", "mpi125pt.py", also without ".py" or "mpi"
/// and with a leading path).  Returns -1 for an unknown name.
inline int bk_stencil_from_file(const char *file) {
  if (!file) return -1;
  const char *base = std::strrchr(file, '/');
  base = base ? base + 1 : file;
  if (std::strncmp(base, "mpi", 3) == 0) base += 3;
  struct Entry {
    const char *name;
    int kind;
  };
  static const Entry table[] = {{"7pt", BK_STAR7},   {"13pt", BK_STAR13},   {"25pt", BK_STAR25},
                                {"125pt", BK_CUBE125}, {"cond", BK_COND7}};
  for (const Entry &e : table) {
    const size_t n = std::strlen(e.name);
    if (std::strncmp(base, e.name, n) == 0 && (base[n] == 0 || std::strcmp(base + n, ".py") == 0))
      return e.kind;
  }
  return -1;
}

/// Element type tag of bElem for the C-ABI.
template <typename T> struct bk_dtype_of;
template <> struct bk_dtype_of<double> { static constexpr int value = BK_F64; };
template <> struct bk_dtype_of<float> { static constexpr int value = BK_F32; };

#endif  // BK_VECSCATTER_H
