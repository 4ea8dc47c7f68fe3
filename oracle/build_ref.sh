#!/bin/bash
# Builds oracle/_ref/libcuttetra.so from the reference's own cuttetraN
# (tetrahedron.f, the subroutine that begins "subroutine cuttetraN(") -- TEST
# INFRASTRUCTURE: the tests use it to pin the oracle's restatement bit for
# bit.  cuttetraN needs only iso_fortran_env and the real_dp macro of
# cppmacro.h, so it compiles on its own; the rest of tetrahedron.f needs the
# reference's modules and is not built.  The subroutine text is read where it
# lies and compiled from a temporary directory outside the repository; only
# the shared library is written under oracle/_ref/.
set -euo pipefail
REF=${SPEX_REFERENCE:-/root/reference}
SRC="$REF/src/tetrahedron.f"
HERE="$(cd "$(dirname "$0")" && pwd)"
FC=${FC_REF:-/opt/rocm/bin/amdflang}
if [ ! -f "$SRC" ] || [ ! -f "$REF/src/cppmacro.h" ]; then
  echo "build_ref: $SRC not present; oracle/_ref not built" >&2
  exit 0
fi
if [ ! -x "$FC" ]; then
  echo "build_ref: $FC not found; oracle/_ref not built" >&2
  exit 0
fi
TMP=$(mktemp -d /tmp/spexref.XXXXXX)
trap 'rm -r "$TMP"' EXIT
{
  echo '# include "cppmacro.h"'
  awk '/^      subroutine cuttetraN\(/{on=1} on{print} on && /^      end *$/{exit}' "$SRC"
} > "$TMP/cuttetran.F"
grep -q "subroutine cuttetraN" "$TMP/cuttetran.F"
mkdir -p "$HERE/_ref"
# macros expanded by the C preprocessor in traditional mode (as the
# reference's own Makefile does for its sources), then fixed-form Fortran
# with long lines
cpp -traditional -P -I"$REF/src" "$TMP/cuttetran.F" > "$TMP/cuttetran.f"
"$FC" -ffixed-line-length-132 -O2 -fPIC -shared -o "$HERE/_ref/libcuttetra.so" "$TMP/cuttetran.f"
