// The device code of libqrdm_amd.so as one translation unit: the kernels of
// panel.hip combine stages defined in kernels.hip (one launch runs several
// steps of a DM iteration); mixed.hip adds the reduced-precision trailing
// update, learned.hip the (tau, delta) network, batched.hip the persistent
// many-matrix kernel, trailing.hip the DM loop's trailing update merged into
// one pass over C per step, wide.hip the DM steps wider than 64 columns.
#include "kernels.hip"
#include "panel.hip"
#include "mixed.hip"
#include "learned.hip"
#include "batched.hip"
#include "trailing.hip"
#include "wide.hip"
