// Horizontal LBP backward launches (paired and unpaired) for 3 label(s) per lane (see
// launch_bwd_pair in lbp_bwd.hip): a translation unit of its own so that the build compiles them
// in parallel.
#define BPL_BWD_PAIR_LPL 3
#include "lbp_bwd.hip"
