#!/bin/sh
# localize_replaced.sh CORRELATION_O [LIBALN_CORR]
#
# Makes local, in the reference's own Correlation.o, every global function that the drop-in library
# (libaln_corr.so) also defines.  A link of the object with the drop-in (as a shared library or as its
# object) then resolves every other object's calls of the replaced functions to the drop-in, and keeps the
# rest of Correlation.o -- the functions listed in kept_from_reference.txt -- from the reference itself.
# (Weakening would not do for the shared library: a weak definition linked into the program still wins
# over a shared library's.)  The object is rewritten in place.
set -e
obj="$1"
lib="${2:-$(dirname "$0")/../lib/libaln_corr.so}"
[ -f "$obj" ] && [ -f "$lib" ] || { echo "usage: $0 Correlation.o [libaln_corr.so]" >&2; exit 2; }
tmp=$(mktemp -d)
trap 'rm -rf "$tmp"' EXIT
nm -g --defined-only "$obj" | awk 'NF == 3 && $2 ~ /[TW]/ {print $3}' | sort -u > "$tmp/obj"
nm -D --defined-only "$lib" | awk 'NF == 3 && $2 ~ /[TW]/ {print $3}' | sort -u > "$tmp/lib"
comm -12 "$tmp/obj" "$tmp/lib" > "$tmp/both"
[ -s "$tmp/both" ] || { echo "no function of $obj is defined by $lib" >&2; exit 1; }
set --
while read -r sym; do set -- "$@" "--localize-symbol=$sym"; done < "$tmp/both"
objcopy "$@" "$obj"
echo "made local $(wc -l < "$tmp/both") functions of $obj"
