// inst_flow.hip - kernel instances for the Henon-Heiles flow under the s = 3 Gauss method (d = 4).
#include "registry.hpp"

namespace cl {

template <class R, int FR = 0>
using S_ = Slot<R, FR>;

typedef Nil F_none;
typedef Cons<S_<QRSpectrum<QR_GIVENS, 4, 4>>, Nil> F_giv4;
typedef Cons<S_<QRSpectrum<QR_HOUSEHOLDER, 4, 4>>, Nil> F_hh4;
typedef Cons<S_<QRSpectrum<QR_MGS, 4, 4>>, Nil> F_mgs4;
typedef Cons<S_<GALI<4, 4>>, Nil> F_gali4;
typedef Cons<S_<MEGNO<4>>, Nil> F_megno4;
typedef Cons<S_<OFLI<4, 4>>, Nil> F_ofli4;
typedef Cons<S_<MEGNO<4>>, Cons<S_<GALI<4, 4>>, Nil>> F_megno_gali4;
// the fused pass of the Henon-Heiles workload: MEGNO (columns 1-2) and GALI_2..GALI_4 on one frame,
// QRSpectrum(GivensQR()) on the second
typedef Cons<S_<MEGNO<4>>, Cons<S_<GALI<4, 4>>, Cons<S_<QRSpectrum<QR_GIVENS, 4, 4>, 1>, Nil>>> F_c4;
// k = 2 and k = 1
typedef Cons<S_<MEGNO<4>>, Nil> F_megno2;
typedef Cons<S_<SALI<4>>, Nil> F_sali2;
typedef Cons<S_<FLI<4>>, Nil> F_fli2;
typedef Cons<S_<GALI<4, 2>>, Nil> F_gali2;
typedef Cons<S_<QRSpectrum<QR_GIVENS, 4, 2>>, Nil> F_giv2;
typedef Cons<S_<MEGNO<4>>, Cons<S_<FLI<4>>, Cons<S_<OFLI<4, 2>>, Cons<S_<SALI<4>>, Nil>>>> F_mfos2;
typedef Cons<S_<Benettin<4>>, Nil> F_ben1;
typedef Cons<S_<Monitor<4, 4>>, Nil> F_mon4;
typedef Cons<S_<QRSpectrum<QR_GIVENS, 4, 4>>, Cons<S_<Monitor<4, 4>>, Nil>> F_giv_mon4;
typedef Cons<S_<OFLI<4, 1>>, Nil> F_ofli1;

#define ADDF(CH, K, NF) if (*n < cap) out[(*n)++] = make_instance<CL_SYS_HENON_HEILES_GAUSS, HHGauss, K, NF, CH>();

// the compensated (double-double) orbit of the element type CL_ELT_DD: state (z, e), s = 8
template <class CH, int K, int NF>
Instance make_dd() {
  Instance in = make_instance<CL_SYS_HENON_HEILES_GAUSS, HHGaussComp, K, NF, CH>();
  in.elt = CL_ELT_DD;
  return in;
}
#define ADDD(CH, K, NF) if (*n < cap) out[(*n)++] = make_dd<CH, K, NF>();

void register_instances_flow(Instance* out, int* n, int cap) {
  ADDD(F_none, 4, 1) ADDD(F_giv4, 4, 1) ADDD(F_megno_gali4, 4, 1) ADDD(F_c4, 4, 2) ADDD(F_giv_mon4, 4, 1)
  ADDD(F_ofli4, 4, 1)
  ADDF(F_none, 4, 1) ADDF(F_giv4, 4, 1) ADDF(F_hh4, 4, 1) ADDF(F_mgs4, 4, 1) ADDF(F_gali4, 4, 1)
  ADDF(F_megno4, 4, 1) ADDF(F_ofli4, 4, 1) ADDF(F_megno_gali4, 4, 1) ADDF(F_c4, 4, 2)
  ADDF(F_mon4, 4, 1) ADDF(F_giv_mon4, 4, 1)
  ADDF(F_none, 2, 1) ADDF(F_megno2, 2, 1) ADDF(F_sali2, 2, 1) ADDF(F_fli2, 2, 1) ADDF(F_gali2, 2, 1)
  ADDF(F_giv2, 2, 1) ADDF(F_mfos2, 2, 1)
  ADDF(F_none, 1, 1) ADDF(F_ben1, 1, 1) ADDF(F_ofli1, 1, 1)
}

}  // namespace cl
