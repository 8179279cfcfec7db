// inst_f32.hip - kernel instances of the FP32-tangent element type (CL_ELT_F64_F32TAN) on the standard-map family:
// FP64 orbit and Jacobian, Float32 frames between handoffs, the reducers in FP64. The chains are those of the
// workloads that run this type: the QR spectra (MGS, Householder), the indicator chains and the fused indicator
// pass with its split passes, and the surrogate's prefix and continuation chains.
#include "registry.hpp"

namespace cl {

template <class R, int FR = 0>
using S_ = Slot<R, FR>;

#define QRF(A, FR) S_<QRSpectrum<A, 2, 2>, FR>
typedef Nil G_none;
typedef Cons<QRF(QR_MGS, 0), Nil> G_mgs;
typedef Cons<QRF(QR_HOUSEHOLDER, 0), Nil> G_hh;
typedef Cons<S_<MEGNO<2>>, Nil> G_megno;
typedef Cons<S_<SALI<2>>, Nil> G_sali;
typedef Cons<S_<FLI<2>>, Nil> G_fli;
typedef Cons<S_<OFLI<2, 2>>, Nil> G_ofli;
typedef Cons<S_<GALI<2, 2>>, Nil> G_gali;
typedef Cons<S_<MEGNO<2>>, Cons<S_<FLI<2>>, Nil>> G_mf;
typedef Cons<S_<MEGNO<2>>, Cons<S_<FLI<2>>, Cons<S_<SALI<2>>, Nil>>> G_mfs;
typedef Cons<S_<MEGNO<2>>, Cons<S_<FLI<2>>, Cons<S_<OFLI<2, 2>>, Cons<S_<SALI<2>>, Nil>>>> G_mfos;
typedef Cons<S_<MEGNO<2>>, Cons<S_<FLI<2>>, Cons<S_<OFLI<2, 2>>, Cons<S_<SALI<2>>, Cons<QRF(QR_HOUSEHOLDER, 1), Nil>>>>> G_mfos_hh;
typedef Cons<S_<MEGNO<2>>, Cons<S_<FLI<2>>, Cons<QRF(QR_MGS, 1), Nil>>> G_prefix_mgs;
typedef Cons<S_<QRSpectrum<QR_MGS, 2, 1>>, Nil> G1_mgs;
typedef Cons<S_<Benettin<2>>, Nil> G1_ben;

#define ADDG(SK, SYS, CH, K, NF) if (*n < cap) out[(*n)++] = make_instance<SK, SYS, K, NF, CH, true>();
#define F32_FOR(SK, SYS)                                                                                      \
  ADDG(SK, SYS, G_none, 2, 1) ADDG(SK, SYS, G_mgs, 2, 1) ADDG(SK, SYS, G_hh, 2, 1) ADDG(SK, SYS, G_megno, 2, 1)  \
  ADDG(SK, SYS, G_sali, 2, 1) ADDG(SK, SYS, G_fli, 2, 1) ADDG(SK, SYS, G_ofli, 2, 1) ADDG(SK, SYS, G_gali, 2, 1) \
  ADDG(SK, SYS, G_mf, 2, 1) ADDG(SK, SYS, G_mfs, 2, 1) ADDG(SK, SYS, G_mfos, 2, 1) ADDG(SK, SYS, G_mfos_hh, 2, 2) \
  ADDG(SK, SYS, G_prefix_mgs, 2, 2) ADDG(SK, SYS, G1_mgs, 1, 1) ADDG(SK, SYS, G1_ben, 1, 1) ADDG(SK, SYS, G_none, 1, 1)

void register_instances_f32(Instance* out, int* n, int cap) {
  F32_FOR(CL_SYS_STANDARD_MAP, StdMap)
  F32_FOR(CL_SYS_WRAPPED_STANDARD_MAP, WrappedStdMap)
  F32_FOR(CL_SYS_STANDARD_MAP_K, StdMapK)
}

}  // namespace cl
