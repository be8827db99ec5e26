// Prints the C4 storage pitches the library uses (c4_pitch, c4_elems in
// vlp4d_internal.hpp) for the grid given on the command line:
//   c4_layout_probe N0 N1 N2 N3  ->  "s1 s2 s3 elems"
// Host code only; built and run by tests/test_c4_layout.py.
#include <cstdio>
#include <cstdlib>

#include "vlp4d_internal.hpp"

int main(int argc, char **argv) {
  if (argc != 5) return 2;
  const long n0 = atol(argv[1]), n1 = atol(argv[2]), n2 = atol(argv[3]), n3 = atol(argv[4]);
  const vlp::C4Pitch p = vlp::c4_pitch(n0, n1, n2);
  printf("%ld %ld %ld %ld\n", p.s1, p.s2, p.s3, vlp::c4_elems(n0, n1, n2, n3));
  return 0;
}
