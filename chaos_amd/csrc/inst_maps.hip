// inst_maps.hip - kernel instances for the two-dimensional maps (standard map, wrapped standard map).
#include "registry.hpp"

namespace cl {

template <class R, int FR = 0>
using S_ = Slot<R, FR>;

// k = 2
typedef Nil C_none;
#define QR2(A, FR) S_<QRSpectrum<A, 2, 2>, FR>
#define QR1(A) S_<QRSpectrum<A, 2, 1>, 0>
typedef Cons<QR2(QR_CGS, 0), Nil> C_cgs;
typedef Cons<QR2(QR_MGS, 0), Nil> C_mgs;
typedef Cons<QR2(QR_RGS, 0), Nil> C_rgs;
typedef Cons<QR2(QR_HOUSEHOLDER, 0), Nil> C_hh;
typedef Cons<QR2(QR_AUGMENTED, 0), Nil> C_aug;
typedef Cons<QR2(QR_GIVENS, 0), Nil> C_giv;
typedef Cons<S_<MEGNO<2>>, Nil> C_megno;
typedef Cons<S_<SALI<2>>, Nil> C_sali;
typedef Cons<S_<FLI<2>>, Nil> C_fli;
typedef Cons<S_<OFLI<2, 2>>, Nil> C_ofli;
typedef Cons<S_<GALI<2, 2>>, Nil> C_gali;
typedef Cons<S_<MEGNO<2>>, Cons<S_<SALI<2>>, Nil>> C_megno_sali;
typedef Cons<S_<FLI<2>>, Cons<S_<SALI<2>>, Nil>> C_fli_sali;
typedef Cons<S_<MEGNO<2>>, Cons<S_<FLI<2>>, Cons<S_<SALI<2>>, Nil>>> C_mfs;
typedef Cons<S_<MEGNO<2>>, Cons<S_<FLI<2>>, Cons<S_<OFLI<2, 2>>, Cons<S_<SALI<2>>, Nil>>>> C_mfos;
// the fused indicator pass: indicators on the first frame, a QR spectrum on the second
typedef Cons<S_<MEGNO<2>>, Cons<S_<FLI<2>>, Cons<S_<OFLI<2, 2>>, Cons<S_<SALI<2>>, Cons<QR2(QR_HOUSEHOLDER, 1), Nil>>>>> C_mfos_hh;
typedef Cons<S_<MEGNO<2>>, Cons<S_<FLI<2>>, Cons<S_<OFLI<2, 2>>, Cons<S_<SALI<2>>, Cons<QR2(QR_MGS, 1), Nil>>>>> C_mfos_mgs;
typedef Cons<S_<MEGNO<2>>, Cons<S_<FLI<2>>, Cons<S_<SALI<2>>, Cons<QR2(QR_HOUSEHOLDER, 1), Nil>>>> C_mfs_hh;
// k = 1
typedef Cons<S_<Benettin<2>>, Nil> C_ben;
typedef Cons<QR1(QR_CGS), Nil> C1_cgs;
typedef Cons<QR1(QR_MGS), Nil> C1_mgs;
typedef Cons<QR1(QR_RGS), Nil> C1_rgs;
typedef Cons<QR1(QR_HOUSEHOLDER), Nil> C1_hh;
typedef Cons<QR1(QR_AUGMENTED), Nil> C1_aug;
typedef Cons<QR1(QR_GIVENS), Nil> C1_giv;
typedef Cons<S_<OFLI<2, 1>>, Nil> C1_ofli;
// ResidualMonitor alone and after the QR spectrum (it then checks the orthonormalised frame)
typedef Cons<S_<Monitor<2, 2>>, Nil> C_mon;
typedef Cons<QR2(QR_MGS, 0), Cons<S_<Monitor<2, 2>>, Nil>> C_mgs_mon;
typedef Cons<S_<Monitor<2, 1>>, Nil> C1_mon;
// the recorder's chains (an estimator, then the non-raising RunMonitor): methods L3, L4, L4b, L5, L5b (k = 2) and
// L1 (Benettin, k = 1) (src/Experiments/run_record.jl:49-53)
typedef Cons<QR2(QR_CGS, 0), Cons<S_<RunMon<2>>, Nil>> C_rec_cgs;
typedef Cons<QR2(QR_MGS, 0), Cons<S_<RunMon<2>>, Nil>> C_rec_mgs;
typedef Cons<QR2(QR_RGS, 0), Cons<S_<RunMon<2>>, Nil>> C_rec_rgs;
typedef Cons<QR2(QR_HOUSEHOLDER, 0), Cons<S_<RunMon<2>>, Nil>> C_rec_hh;
typedef Cons<QR2(QR_AUGMENTED, 0), Cons<S_<RunMon<2>>, Nil>> C_rec_aug;
typedef Cons<S_<Benettin<2>>, Cons<S_<RunMon<2>>, Nil>> C1_rec_ben;
// StretchingNumbers (a workspace reducer, k = 2) alone and as the indicator recorder's chain with RunMonitor
typedef Cons<S_<Stretch<2>>, Nil> C_st;
typedef Cons<S_<Stretch<2>>, Cons<S_<RunMon<2>>, Nil>> C_st_mon;
// the indicator recorder's tangent kind: members then RunMonitor (src/Experiments/run_indicator_chain.jl:27-31; the
// study's spec is FLI, MEGNO, SALI), and each member alone
typedef Cons<S_<FLI<2>>, Cons<S_<MEGNO<2>>, Cons<S_<SALI<2>>, Cons<S_<RunMon<2>>, Nil>>>> C_ind_fms;
typedef Cons<S_<MEGNO<2>>, Cons<S_<FLI<2>>, Cons<S_<SALI<2>>, Cons<S_<RunMon<2>>, Nil>>>> C_ind_mfs;
typedef Cons<S_<FLI<2>>, Cons<S_<RunMon<2>>, Nil>> C_ind_f;
typedef Cons<S_<MEGNO<2>>, Cons<S_<RunMon<2>>, Nil>> C_ind_m;
typedef Cons<S_<SALI<2>>, Cons<S_<RunMon<2>>, Nil>> C_ind_s;
// WeightedBirkhoff (k = 1) alone and after RunMonitor (the indicator recorder's window-kind order)
typedef Cons<S_<WBirk>, Nil> C1_wb;
typedef Cons<S_<RunMon<2>>, Cons<S_<WBirk>, Nil>> C1_mon_wb;
// WeightedPhaseAdvance (k = 1, one or two windows) alone and after RunMonitor (the phase-advance window kind)
typedef Cons<S_<WPA<1>>, Nil> C1_pa1;
typedef Cons<S_<WPA<2>>, Nil> C1_pa2;
typedef Cons<S_<RunMon<2>>, Cons<S_<WPA<1>>, Nil>> C1_mon_pa1;
typedef Cons<S_<RunMon<2>>, Cons<S_<WPA<2>>, Nil>> C1_mon_pa2;

#define ADD2(SK, SYS, CH, NF) if (*n < cap) out[(*n)++] = make_instance<SK, SYS, 2, NF, CH>();
#define ADD1(SK, SYS, CH) if (*n < cap) out[(*n)++] = make_instance<SK, SYS, 1, 1, CH>();
#define ALL_FOR(SK, SYS)                                                                     \
  ADD2(SK, SYS, C_none, 1) ADD2(SK, SYS, C_cgs, 1) ADD2(SK, SYS, C_mgs, 1) ADD2(SK, SYS, C_rgs, 1) \
  ADD2(SK, SYS, C_hh, 1) ADD2(SK, SYS, C_aug, 1) ADD2(SK, SYS, C_giv, 1) ADD2(SK, SYS, C_megno, 1) \
  ADD2(SK, SYS, C_sali, 1) ADD2(SK, SYS, C_fli, 1) ADD2(SK, SYS, C_ofli, 1) ADD2(SK, SYS, C_gali, 1) \
  ADD2(SK, SYS, C_megno_sali, 1) ADD2(SK, SYS, C_fli_sali, 1) ADD2(SK, SYS, C_mfs, 1)           \
  ADD2(SK, SYS, C_mfos, 1) ADD2(SK, SYS, C_mfos_hh, 2) ADD2(SK, SYS, C_mfos_mgs, 2)            \
  ADD2(SK, SYS, C_mfs_hh, 2) ADD2(SK, SYS, C_mon, 1) ADD2(SK, SYS, C_mgs_mon, 1)             \
  ADD1(SK, SYS, C_none) ADD1(SK, SYS, C_ben) ADD1(SK, SYS, C1_cgs) ADD1(SK, SYS, C1_mgs)        \
  ADD1(SK, SYS, C1_rgs) ADD1(SK, SYS, C1_hh) ADD1(SK, SYS, C1_aug) ADD1(SK, SYS, C1_giv)        \
  ADD1(SK, SYS, C1_ofli) ADD1(SK, SYS, C1_mon) ADD2(SK, SYS, C_rec_cgs, 1) ADD2(SK, SYS, C_rec_mgs, 1)          \
  ADD2(SK, SYS, C_rec_rgs, 1) ADD2(SK, SYS, C_rec_hh, 1) ADD2(SK, SYS, C_rec_aug, 1) ADD1(SK, SYS, C1_rec_ben) \
  ADD2(SK, SYS, C_st, 1) ADD2(SK, SYS, C_st_mon, 1) ADD1(SK, SYS, C1_wb) ADD1(SK, SYS, C1_mon_wb)               \
  ADD2(SK, SYS, C_ind_fms, 1) ADD2(SK, SYS, C_ind_mfs, 1) ADD2(SK, SYS, C_ind_f, 1) ADD2(SK, SYS, C_ind_m, 1)   \
  ADD2(SK, SYS, C_ind_s, 1) ADD1(SK, SYS, C1_pa1) ADD1(SK, SYS, C1_pa2) ADD1(SK, SYS, C1_mon_pa1)              \
  ADD1(SK, SYS, C1_mon_pa2)

// the surrogate workload on the per-orbit-K map: the traced prefix (MEGNO + FLI on frame 1, the QR spectrum
// on frame 2), the continuation and the plain method (the QR spectrum alone)
typedef Cons<S_<MEGNO<2>>, Cons<S_<FLI<2>>, Cons<QR2(QR_MGS, 1), Nil>>> C_prefix_mgs;
typedef Cons<S_<MEGNO<2>>, Cons<S_<FLI<2>>, Nil>> C_mf;

// the pair kind of the indicator recorder (run_indicator_chain.jl:27-31, the study's spec RLI, MEGNO) on the
// product map, and RLI alone; RunMonitor reads orbit A's 2 x 2 tangent
typedef Cons<S_<RLI<4>>, Cons<S_<MEGNO<4>>, Cons<S_<RunMon<2>>, Nil>>> C_pair_rm_mon;
typedef Cons<S_<RLI<4>>, Cons<S_<RunMon<2>>, Nil>> C_pair_r_mon;
typedef Cons<S_<RLI<4>>, Nil> C_pair_r;
#define PAIR_FOR(SK, SYS) \
  ADD2(SK, SYS, C_pair_rm_mon, 1) ADD2(SK, SYS, C_pair_r_mon, 1) ADD2(SK, SYS, C_pair_r, 1)

void register_instances_map(Instance* out, int* n, int cap) {
  PAIR_FOR(CL_SYS_PAIR_STANDARD_MAP, PairMap<StdMap>)
  PAIR_FOR(CL_SYS_PAIR_WRAPPED_STANDARD_MAP, PairMap<WrappedStdMap>)
  ALL_FOR(CL_SYS_STANDARD_MAP, StdMap)
  ALL_FOR(CL_SYS_WRAPPED_STANDARD_MAP, WrappedStdMap)
  ADD2(CL_SYS_STANDARD_MAP_K, StdMapK, C_none, 1) ADD2(CL_SYS_STANDARD_MAP_K, StdMapK, C_mgs, 1)
  ADD2(CL_SYS_STANDARD_MAP_K, StdMapK, C_hh, 1) ADD2(CL_SYS_STANDARD_MAP_K, StdMapK, C_mf, 1)
  ADD2(CL_SYS_STANDARD_MAP_K, StdMapK, C_mfos_hh, 2) ADD2(CL_SYS_STANDARD_MAP_K, StdMapK, C_prefix_mgs, 2)
  ADD2(CL_SYS_STANDARD_MAP, StdMap, C_prefix_mgs, 2) ADD2(CL_SYS_STANDARD_MAP, StdMap, C_mf, 1)
}

}  // namespace cl
