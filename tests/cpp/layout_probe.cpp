// Prints the size and field offsets of the device records the invariant
// tests read back, for comparison with their numpy dtypes.
#include <cstddef>
#include <cstdio>

#include "ggks_types.hpp"

using namespace ggks;

#define F(T, f) std::printf(#T " " #f " %zu\n", offsetof(T, f))
#define S(T) std::printf(#T " itemsize %zu\n", sizeof(T))

int main() {
  S(SegDesc); F(SegDesc, data_off); F(SegDesc, len); F(SegDesc, row_base); F(SegDesc, red_base);
  F(SegDesc, lo); F(SegDesc, hi); F(SegDesc, bmax); F(SegDesc, pbits); F(SegDesc, nchunks);
  F(SegDesc, sample_off); F(SegDesc, nsample); F(SegDesc, lut_off); F(SegDesc, tot_off);
  F(SegDesc, bits_off); F(SegDesc, q_off); F(SegDesc, q_cnt); F(SegDesc, mode);
  F(SegDesc, radix_bits); F(SegDesc, qsrc); F(SegDesc, out_off); F(SegDesc, flags);
  S(SegState); F(SegState, kb); F(SegState, kmax); F(SegState, lmax); F(SegState, hv);
  F(SegState, s); F(SegState, P); F(SegState, B); F(SegState, nact); F(SegState, poff);
  F(SegState, pwid); F(SegState, rsh); F(SegState, twid); F(SegState, lsh); F(SegState, hvon);
  F(SegState, hb); F(SegState, nfree); F(SegState, arrive); F(SegState, nred);
  S(Chunk); F(Chunk, seg); F(Chunk, idx); F(Chunk, start); F(Chunk, len);
  S(Child); F(Child, start); F(Child, len); F(Child, lo); F(Child, hi); F(Child, q_off);
  F(Child, q_cnt); F(Child, src); F(Child, need_range); F(Child, bucket); F(Child, seg);
  S(Query); F(Query, rank); F(Query, out);
  return 0;
}
