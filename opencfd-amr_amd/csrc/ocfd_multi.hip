// Multispecies (CHEM) flow path on the device: EBR::compute_dSdt_multi on
// regular boxes (Source/CHEM/FlowAdvance.cpp:92-496), the stage updates of
// EBR::flow_advance_multi (:9-90) and the Strang-split EBR::advance
// (Source/Advance.cpp:152-177).  FP64.
//
// Per RK stage:
//   k_mprims     on grow(bx,3): c2prim_rgas (Newton temperature, CKPY,
//                gamma = 1 + p/e) and, viscous, the mixture viscosity,
//                conductivity and diffusion coefficients (19 comps)
//   k_mface<D>   per face: Steger-Warming split of the flow (cell gamma) and
//                of the 9 partial densities at the 6 stencil cells, WENO-Z5
//                for 14 comps, Fickian diffusion and the multispecies
//                viscous flux (ChemViscous.H)
//   k_mdiv       per valid cell: divergence of the 14 face fluxes and the
//                stage combination of state and species
#include <map>
#include <mutex>
#include <string>

#include "ocfd_device.hpp"
#include "ocfd_internal.hpp"
#include "ocfd_chem_device.hpp"

namespace ocfd {

using chem::NS;

namespace tr {

// transport fits of Source/Constants.cpp:84-125: sqrt(mu/sqrt(T)),
// lambda/sqrt(T) and D_jk/T^1.5 as quartics in ln T
__device__ constexpr double visc_fit[NS][5] = {
    {-1.85174759e-04, 3.89197447e-04, -6.98160471e-05, 6.36454790e-06, -2.11216442e-07},
    {-5.80149070e-03, 3.39295201e-03, -6.36767978e-04, 5.43779133e-05, -1.73237416e-06},
    {1.49332282e-02, -8.19367041e-03, 1.68282727e-03, -1.45738528e-04, 4.60116686e-06},
    {-5.03440195e-03, 2.77694310e-03, -5.05779281e-04, 4.17970313e-05, -1.28989348e-06},
    {-4.25439228e-03, 2.69409815e-03, -5.12461257e-04, 4.46919118e-05, -1.45052986e-06},
    {-4.31987570e-03, 2.73556559e-03, -5.20349037e-04, 4.53798077e-05, -1.47285635e-06},
    {-5.84665026e-03, 3.41936319e-03, -6.41724663e-04, 5.48011981e-05, -1.74585918e-06},
    {-5.89078698e-03, 3.44517618e-03, -6.46569083e-04, 5.52148957e-05, -1.75903878e-06},
    {-4.86883676e-03, 2.91081563e-03, -5.49840279e-04, 4.73395563e-05, -1.51957597e-06},
};
__device__ constexpr double cond_fit[NS][5] = {
    {-5.28463600e-01, 3.15473947e-01, -6.87861093e-02, 6.58583530e-03, -2.30878456e-04},
    {3.45755008e-02, -2.10056875e-02, 4.78193956e-03, -4.69828968e-04, 1.73609698e-05},
    {-2.42370932e-01, 1.55066730e-01, -3.67910197e-02, 3.82882196e-03, -1.45626051e-04},
    {-2.05169993e-01, 1.03579992e-01, -1.83028199e-02, 1.47068561e-03, -4.35660111e-05},
    {5.22190001e-02, -3.07414624e-02, 7.02757543e-03, -7.03400778e-04, 2.64083229e-05},
    {-1.41026759e-01, 8.85651492e-02, -2.02651378e-02, 2.03706447e-03, -7.50266004e-05},
    {-7.12040025e-02, 4.31981643e-02, -9.82859052e-03, 1.00407694e-03, -3.78442817e-05},
    {-1.18157269e-02, 9.27318610e-03, -2.71924641e-03, 3.59761794e-04, -1.62746865e-05},
    {4.77081884e-02, -2.51859514e-02, 4.94733498e-03, -4.15782937e-04, 1.29219096e-05},
};
__device__ constexpr double diff_fit[45][5] = {
    {-0.006098424587788583, 0.004074726320140367, -0.0007655079217743792, 7.01945414920243e-05, -2.283978048241683e-06},
    {-0.007361772345996411, 0.0042187108652972285, -0.0007832424157195558, 6.805076239783791e-05, -2.1468141590124147e-06},
    {-0.012714368881604958, 0.005800464307529211, -0.0008628173792959883, 5.797865818973548e-05, -1.3148664989894858e-06},
    {-0.02331895278393948, 0.012988306044054893, -0.0023909974764966175, 0.00020481645048811995, -6.393432695483336e-06},
    {-0.008203316666546017, 0.004854681011212246, -0.0009069544030339767, 7.98972126021157e-05, -2.5440602562722086e-06},
    {-0.008176066416389936, 0.004838554451993892, -0.0009039416296196897, 7.963180543591231e-05, -2.535609250271386e-06},
    {-0.007355106588534061, 0.004214891010226138, -0.000782533224544873, 6.798914545886926e-05, -2.144870314265137e-06},
    {-0.00734883038938788, 0.0042112943940458315, -0.0007818654797209853, 6.793112951970908e-05, -2.143040071146875e-06},
    {-0.006631024777860074, 0.0038397370846872426, -0.0007137696048348218, 6.226499117199988e-05, -1.968810850179303e-06},
    {-0.003165622529780972, 0.0016565295961648823, -0.0002955102867456441, 2.4323268960998937e-05, -7.332687938295335e-07},
    {0.0017962300955720687, -0.0016252040565658543, 0.00047039226523034656, -5.1041731991760684e-05, 1.9707403545544835e-06},
    {-0.02065558573348347, 0.01049932817565325, -0.0018228205967741945, 0.00014586133392843038, -4.265266434634806e-06},
    {-0.004477615809775433, 0.002401743913536847, -0.00043555046112257176, 3.6523201605646507e-05, -1.1207561496132533e-06},
    {-0.004388261916818509, 0.002353815467310766, -0.00042685875309455987, 3.579435608041233e-05, -1.0983906923530078e-06},
    {-0.003141360101617588, 0.0016438333792438298, -0.0002932453934942289, 2.4136846998057358e-05, -7.276487676670712e-07},
    {-0.0031183633828342915, 0.001631799491779716, -0.00029109865398359785, 2.3960150196485035e-05, -7.223219240282344e-07},
    {-0.0030766151054801255, 0.001623930137648841, -0.0002916902121621802, 2.4185952795909278e-05, -7.345781692660118e-07},
    {0.015418678757875483, -0.008276025270343653, 0.0016022988850484384, -0.0001292488075144311, 3.7780416116707573e-06},
    {0.015491646067634311, -0.012297762819716031, 0.0033220236270480115, -0.00034981378874120536, 1.3233960398093872e-05},
    {-0.0031409034359764906, 0.000964215499007546, -1.5256977763517157e-05, -1.0009613901179592e-05, 6.865263141553794e-07},
    {-0.00309121182555696, 0.0009489608368058763, -1.501560014564807e-05, -9.851253784321545e-06, 6.756649174673959e-07},
    {-9.806367280891379e-05, -0.0005259982739909082, 0.00023777270488750687, -2.9610758656448636e-05, 1.2409556192468912e-06},
    {-9.754926582674518e-05, -0.0005232390750052945, 0.00023652543424357053, -2.945543119768198e-05, 1.2344460095131642e-06},
    {0.0009009636432407803, -0.0010975268417085945, 0.0003568805024623778, -4.047906132942509e-05, 1.6078992712041145e-06},
    {-0.049547371639744826, 0.02420786673131437, -0.0040039853524907396, 0.0003046592732139884, -8.373772281551857e-06},
    {-0.027018754679375097, 0.014131423473429378, -0.0025198932193017465, 0.00020732039220446453, -6.247265305088386e-06},
    {-0.026971255852400414, 0.01410658050614936, -0.002515463260429814, 0.00020695592405809058, -6.236282645924077e-06},
    {-0.020645950902716106, 0.01049443074735431, -0.0018219703392133265, 0.00014579329667774188, -4.2632768943398505e-06},
    {-0.02063688301930936, 0.010489821500976207, -0.0018211701137997654, 0.00014572926297830028, -4.261404425595428e-06},
    {-0.019430526862365516, 0.00997359743800644, -0.0017481642351558765, 0.00014128443680858248, -4.177030598880565e-06},
    {-0.0059382904901673595, 0.0032649962117391896, -0.0005991311922260038, 5.1014036740744293e-05, -1.5859623241784843e-06},
    {-0.00584963856139101, 0.0032162535286223044, -0.0005901868443753459, 5.02524551442533e-05, -1.5622857089560921e-06},
    {-0.0044547665185479274, 0.0023894878048250404, -0.00043332784539284135, 3.633682356306633e-05, -1.1150369265388267e-06},
    {-0.004433163993709295, 0.0023779004479025216, -0.00043122650618556915, 3.616061519603177e-05, -1.109629771571358e-06},
    {-0.00419972224671333, 0.0022703026942310655, -0.0004131673665395828, 3.48089622951453e-05, -1.0723125190961097e-06},
    {-0.005759622266467147, 0.003166760688484433, -0.0005811048417034481, 4.947915269564291e-05, -1.538244707164061e-06},
    {-0.004364944907483492, 0.002341308479746039, -0.0004245906410903713, 3.5604162935439344e-05, -1.0925543985047233e-06},
    {-0.004342895613050795, 0.0023294814805234227, -0.0004224458433307047, 3.542431033976762e-05, -1.0870354162206174e-06},
    {-0.004119743857368838, 0.0022270676557828092, -0.0004052991175070151, 3.414606970952483e-05, -1.05189168573903e-06},
    {-0.003116908817255142, 0.0016310383363007772, -0.00029096287055129454, 2.3948973946122656e-05, -7.219849958113728e-07},
    {-0.0030937303493271603, 0.0016189093418438713, -0.00028879916479071873, 2.37708806629901e-05, -7.166160527175076e-07},
    {-0.0030546059172749706, 0.0016123130250736487, -0.00028960354725398777, 2.40129336925455e-05, -7.293232158117316e-07},
    {-0.003070376910660832, 0.001606688787447522, -0.00028661912554336986, 2.35914429806519e-05, -7.112065802854437e-07},
    {-0.0030337554543677816, 0.0016013075226178426, -0.00028762673971045356, 2.3849023585380798e-05, -7.243449216977582e-07},
    {-0.0029800170761932095, 0.0015859694924580709, -0.000286608674495536, 2.3921390316404743e-05, -7.312450753144484e-07},
};


#pragma clang fp contract(off)

__device__ __forceinline__ double dot5(const double* x, const double* y)
{
    return x[0] * y[0] + x[1] * y[1] + x[2] * y[2] + x[3] * y[3] + x[4] * y[4];
}

// mixtureProperties, Source/CHEM/ChemViscous.H:16-64 (Wilke's rule)
__device__ __forceinline__ void mixture(double T, const double* X, double& mu, double& lambda)
{
    double mui[NS], sqmui[NS], lambdai[NS];
    const double sqT = sqrt(T), logT = log(T);
    const double sqsqT = sqrt(sqT);
    const double poly[5] = {1.0, logT, logT * logT, logT * logT * logT, logT * logT * logT * logT};
#pragma unroll
    for (int k = 0; k < NS; k++) {
        sqmui[k] = sqsqT * dot5(poly, visc_fit[k]);
        lambdai[k] = sqT * dot5(poly, cond_fit[k]);
        mui[k] = (sqmui[k] * sqmui[k]);
    }
    double sum1 = 0.0, sum2 = 0.0;
#pragma unroll
    for (int k = 0; k < NS; k++) {
        sum1 += X[k] * lambdai[k];
        sum2 += X[k] / lambdai[k];
    }
    lambda = 0.5 * (sum1 + 1.0 / sum2);
    // Wilke's phi_kj = (1 + sqrt(mu_k/mu_j) (M_j/M_k)^(1/4))^2 / sqrt(8 (1 + M_k/M_j))
    // and phi_jk = phi_kj (mu_j/mu_k)(M_k/M_j): the reference's formulas with
    // one reciprocal per species root instead of three divisions per pair
    // (the molecular-weight factors are compile-time constants); the values
    // agree with the reference order to a few ulps
    double rsq[NS];
#pragma unroll
    for (int k = 0; k < NS; k++) rsq[k] = 1.0 / sqmui[k];
    double phi[NS][NS];
#pragma unroll
    for (int k = 0; k < NS; k++) {
#pragma unroll
        for (int j = 0; j < k; j++) {
            const double wratiojk = chem::mw[j] / chem::mw[k];
            const double factor1 = 1.0 + (sqmui[k] * rsq[j]) * sqrt(sqrt(wratiojk));
            const double tmp = factor1 * factor1 * (1.0 / sqrt(8.0 + 8.0 * chem::mw[k] / chem::mw[j]));
            const double rjk = sqmui[j] * rsq[k];
            phi[k][j] = tmp;
            phi[j][k] = tmp * (rjk * rjk) * (chem::mw[k] / chem::mw[j]);
        }
        phi[k][k] = 1.0;
    }
    mu = 0;
#pragma unroll
    for (int k = 0; k < NS; k++) {
        double tmp = 0;
#pragma unroll
        for (int j = 0; j < NS; j++) tmp += X[j] * phi[k][j];
        mu += X[k] * mui[k] / tmp;
    }
}

// getMixDiffCoeffsMass, Source/CHEM/ChemViscous.H:66-106
__device__ __forceinline__ void diff_coeffs(double T, double p, const double* X, double* D)
{
    const double logT = log(T);
    const double poly[5] = {1.0, logT, logT * logT, logT * logT * logT, logT * logT * logT * logT};
    const double t15 = T * sqrt(T);
    double mmw = 0;
#pragma unroll
    for (int n = 0; n < NS; ++n) mmw += X[n] * chem::mw[n];
    // reciprocals of the binary coefficients of the upper triangle, row by
    // row (the fit order): each is divided by twice in the sums below
    double rbd[45];
#pragma unroll
    for (int ic = 0; ic < 45; ++ic) rbd[ic] = 1.0 / (t15 * dot5(poly, diff_fit[ic]));
#pragma unroll
    for (int n = 0; n < NS; n++) {
        double sum1 = 0.0, sum2 = 0.0;
#pragma unroll
        for (int nn = 0; nn < NS; nn++) {
            if (nn == n) continue;
            const int a = n < nn ? n : nn, b = n < nn ? nn : n;
            const int ic = a * NS - a * (a - 1) / 2 + (b - a);   // index of (a, b) in the upper triangle
            sum1 += X[nn] * rbd[ic];
            sum2 += X[nn] * chem::mw[nn] * rbd[ic];
        }
        sum1 *= p;
        sum2 *= p * X[n] / (mmw - chem::mw[n] * X[n]);
        D[n] = 1.0 / (sum1 + sum2);
    }
}

// sensible enthalpy / R in K, Source/CHEM/LiDryer.H:616-695
__device__ __forceinline__ void species_enthalpy(double T, double* h)
{
    const double tc[5] = {T, T * T, T * T * T, T * T * T * T, T * T * T * T * T};
    const bool low = T < 1000;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        const double* a = low ? chem::lo[s] : chem::hi[s];
        double v = a[0] * tc[0] + a[1] / 2.0 * tc[1] + a[2] / 3.0 * tc[2] + a[3] / 4.0 * tc[3] + a[4] / 5.0 * tc[4];
        if (!low) v = v + (chem::hi[s][5] - chem::lo[s][5]);
        h[s] = v;
    }
}

#pragma clang fp contract(on)

} // namespace tr

enum { MQ_R = 0, MQ_U, MQ_V, MQ_W, MQ_P, MQ_T, MQ_C, MQ_G, MQ_MU, MQ_LAM, MQ_D, MQ_N = MQ_D + NS };
constexpr int NF = NC + NS;   // flow + species face flux comps

struct MultiArgs {
    StageArgs a;              // flow state, coefficients, flags, control
    const double* pcur;       // species of the current stage (4 ghosts, 9 comps)
    const double* pold;       // species U^n, or null
    double* pnew;             // species stage output, or null
    double* pacc;             // species dSdt of a plain compute (valid layout, 9 comps)
    double* mq;               // 19-comp primitive / transport scratch, grown layout
    double* mF[3];            // face fluxes per direction, 14 comps, grown layout
};

__device__ __forceinline__ long gl(const BoxDesc& B, int i, int j, int k)
{
    return ((long)(k - B.lo[2] + NG) * B.nyg + (j - B.lo[1] + NG)) * B.nxg + (i - B.lo[0] + NG);
}

#pragma clang fp contract(off)

// c2prim_rgas (Source/kernel/Kernels.H:42-81) and the transport properties
// (FlowAdvance.cpp:188-202) on grow(bx,3) of a regular box
__global__ __launch_bounds__(256) void k_mprims(const BoxDesc* __restrict__ boxes, const int4* __restrict__ tiles,
                                               MultiArgs m)
{
    if (ctl_done(m.a.ctl)) return;
    const int4 t = tiles[blockIdx.x];
    for (int c = threadIdx.x; c < t.z; c += 256) {   // tiles hold at most 256 cells
        const BoxDesc B = boxes[t.x];
        const int g = t.w;
        const int nx = B.hi[0] - B.lo[0] + 1 + 2 * g, ny = B.hi[1] - B.lo[1] + 1 + 2 * g;
        const int cc = t.y + c;
        const int i = B.lo[0] - g + cc % nx, j = B.lo[1] - g + (cc / nx) % ny, k = B.lo[2] - g + cc / (nx * ny);
        const long l0 = gl(B, i, j, k), st = B.nptg;
        const double* S = m.a.scur + B.off_s + l0;
        const double* P = m.pcur + B.off_s / NC * NS + l0;
        const double rho = fmax(S[0], m.a.ph.smallr);
        const double rhoinv = 1.0 / rho;
        const double ux = S[st] * rhoinv, uy = S[2 * st] * rhoinv, uz = S[3 * st] * rhoinv;
        const double kineng = 0.5 * (ux * ux + uy * uy + uz * uz);
        const double ei = S[4 * st] - rho * kineng;
        double rhoi[NS];
#pragma unroll
        for (int n = 0; n < NS; ++n) rhoi[n] = P[n * st];
        const double T = chem::T_given_ey(ei, rhoi);
        // (divisions by the constant molecular weights and by rho as
        // multiplications by their reciprocals: a last-bit difference at most)
        double yow = 0;
#pragma unroll
        for (int n = 0; n < NS; ++n) yow += rhoi[n] * (1.0 / chem::mw[n]);
        const double p = chem::Ru * T * yow;
        const double gama = p / ei + 1.0;
        double* q = m.mq + B.off_s / NC * MQ_N + l0;
        q[MQ_R * st] = rho; q[MQ_U * st] = ux; q[MQ_V * st] = uy; q[MQ_W * st] = uz;
        q[MQ_P * st] = p; q[MQ_T * st] = T; q[MQ_C * st] = sqrt(gama * p / rho); q[MQ_G * st] = gama;
        // the transport properties are read at the two cells of each face,
        // all within grow(bx,1)
        const bool face_cell = i >= B.lo[0] - 1 && i <= B.hi[0] + 1 && j >= B.lo[1] - 1 && j <= B.hi[1] + 1 &&
                               k >= B.lo[2] - 1 && k <= B.hi[2] + 1;
        if ((m.a.flags & OCFD_VISC) && face_cell) {
            double Y[NS], X[NS];
#pragma unroll
            for (int n = 0; n < NS; ++n) Y[n] = rhoi[n] * rhoinv;
            // CKYTX, LiDryer.H:36-52
            double tmp[NS], yw = 0;
#pragma unroll
            for (int n = 0; n < NS; ++n) tmp[n] = Y[n] * (1.0 / chem::mw[n]);
#pragma unroll
            for (int n = 0; n < NS; ++n) yw += tmp[n];
            const double ywinv = 1.0 / yw;
#pragma unroll
            for (int n = 0; n < NS; ++n) X[n] = tmp[n] * ywinv;
            double mu, lam, D[NS];
            tr::mixture(T, X, mu, lam);
            tr::diff_coeffs(T, p, X, D);
            q[MQ_MU * st] = mu;
            q[MQ_LAM * st] = lam;
#pragma unroll
            for (int n = 0; n < NS; ++n) q[(MQ_D + n) * st] = D[n];
        }
    }
}

// Steger-Warming split of the flow at one cell (FluxSplit.H:11-193 with the
// cell's gamma) and the factors of the species split (FlowAdvance.cpp:214-242)
template <int D>
__device__ __forceinline__ void msplit(const double* q, long st, double* fp, double* fm, double& sp, double& sm)
{
    const double d = q[MQ_R * st], u = q[MQ_U * st], v = q[MQ_V * st], w = q[MQ_W * st], c = q[MQ_C * st];
    const double gamma = q[MQ_G * st];
    const double tmp1 = 1.0 / (2.0 * gamma);
    const double tmp2 = 2.0 * (gamma - 1.0);
    const double tmp3 = (3.0 - gamma) / (2.0 * (gamma - 1.0));
    const double un = D == 0 ? u : (D == 1 ? v : w);
    const double uc1 = un - c, uc2 = un + c;
    double vvc1, vvc2;
    if (D == 0) { vvc1 = 0.5 * (uc1 * uc1 + v * v + w * w); vvc2 = 0.5 * (uc2 * uc2 + v * v + w * w); }
    else if (D == 1) { vvc1 = 0.5 * (u * u + uc1 * uc1 + w * w); vvc2 = 0.5 * (u * u + uc2 * uc2 + w * w); }
    else { vvc1 = 0.5 * (u * u + v * v + uc1 * uc1); vvc2 = 0.5 * (u * u + v * v + uc2 * uc2); }
    const double v2 = (gamma - 1) * (u * u + v * v + w * w);
    const double w2 = tmp3 * c * c;
    const double E1 = un, E2 = un - c, E3 = un + c;
    const double E1P = (E1 + fabs(E1)) * 0.5, E2P = (E2 + fabs(E2)) * 0.5, E3P = (E3 + fabs(E3)) * 0.5;
    const double E1M = E1 - E1P, E2M = E2 - E2P, E3M = E3 - E3P;
    const double tmp0 = d * tmp1;
    const double vel[3] = {u, v, w};
    fp[0] = tmp0 * (tmp2 * E1P + E2P + E3P);
    fm[0] = tmp0 * (tmp2 * E1M + E2M + E3M);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        if (a == D) {
            fp[1 + a] = tmp0 * (tmp2 * E1P * un + E2P * uc1 + E3P * uc2);
            fm[1 + a] = tmp0 * (tmp2 * E1M * un + E2M * uc1 + E3M * uc2);
        } else {
            fp[1 + a] = tmp0 * vel[a] * (tmp2 * E1P + E2P + E3P);
            fm[1 + a] = tmp0 * vel[a] * (tmp2 * E1M + E2M + E3M);
        }
    }
    fp[4] = tmp0 * (E1P * v2 + E2P * vvc1 + E3P * vvc2 + w2 * (E2P + E3P));
    fm[4] = tmp0 * (E1M * v2 + E2M * vvc1 + E3M * vvc2 + w2 * (E2M + E3M));
    // species flux factors: f = (rho_k tmp1) (tmp2 E1 + E2 + E3)
    sp = tmp2 * E1P + E2P + E3P;
    sm = tmp2 * E1M + E2M + E3M;
}

#pragma clang fp contract(on)

template <int D>
__global__ __launch_bounds__(256) void k_mface(const BoxDesc* __restrict__ boxes, const int4* __restrict__ tiles,
                                              MultiArgs m)
{
    if (ctl_done(m.a.ctl)) return;
    const int4 t = tiles[blockIdx.x];
    for (int c = threadIdx.x; c < t.z; c += 256) {   // tiles hold at most 256 cells
        const BoxDesc B = boxes[t.x];
        const int di = D == 0, dj = D == 1, dk = D == 2;
        const int nx = B.hi[0] - B.lo[0] + 1 + di, ny = B.hi[1] - B.lo[1] + 1 + dj;
        const int cc = t.y + c;
        const int i = B.lo[0] + cc % nx, j = B.lo[1] + (cc / nx) % ny, k = B.lo[2] + cc / (nx * ny);
        const long st = B.nptg;
        const long ds = D == 0 ? 1 : (D == 1 ? (long)B.nxg : (long)B.nxg * B.nyg);
        const long l0 = gl(B, i, j, k);
        const double* q = m.mq + B.off_s / NC * MQ_N;
        const double* P = m.pcur + B.off_s / NC * NS;
        // split fluxes at cells i-3 .. i+2 (index s = 0..5)
        double fp[6][NC], fm[6][NC], sp[6], sm[6], t1[6];
#pragma unroll
        for (int s = 0; s < 6; ++s) {
            const long l = l0 + (s - 3) * ds;
            msplit<D>(q + l, st, fp[s], fm[s], sp[s], sm[s]);
            t1[s] = 1.0 / (2.0 * q[MQ_G * st + l]);
        }
        double F[NF];
#pragma unroll
        for (int n = 0; n < NC; ++n)
            F[n] = weno_p<double>(fp[0][n], fp[1][n], fp[2][n], fp[3][n], fp[4][n]) +
                   weno_m<double>(fm[1][n], fm[2][n], fm[3][n], fm[4][n], fm[5][n]);
#pragma unroll
        for (int n = 0; n < NS; ++n) {
            double a[6], b[6];
#pragma unroll
            for (int s = 0; s < 6; ++s) {
                const double tmp0 = P[n * st + l0 + (s - 3) * ds] * t1[s];
                a[s] = tmp0 * sp[s];
                b[s] = tmp0 * sm[s];
            }
            F[NC + n] = weno_p<double>(a[0], a[1], a[2], a[3], a[4]) + weno_m<double>(b[1], b[2], b[3], b[4], b[5]);
        }
        if (m.a.flags & OCFD_VISC) {
            const long lm = l0 - ds;
            const double dxn = B.dxinv[D];
            // diffusion_{x,y,z}, ChemViscous.H:108-199
            const double rhor = q[MQ_R * st + l0], rhol = q[MQ_R * st + lm];
            const double rho = 0.5 * (rhol + rhor);
            double J[NS];
#pragma unroll
            for (int n = 0; n < NS; ++n) {
                const double Ytr = P[n * st + l0] / rhor;
                const double Ytl = P[n * st + lm] / rhol;
                const double dn = 0.5 * (q[(MQ_D + n) * st + l0] + q[(MQ_D + n) * st + lm]);
                J[n] = rho * dn * (Ytl - Ytr) * dxn;
                F[NC + n] += J[n];
            }
            // compute_visc_{x,y,z}_multi, ChemViscous.H:201-357
            const double Tr = q[MQ_T * st + l0], Tl = q[MQ_T * st + lm];
            const double T = 0.5 * (Tl + Tr);
            const double muf = 0.5 * (q[MQ_MU * st + l0] + q[MQ_MU * st + lm]);
            const double lamf = 0.5 * (q[MQ_LAM * st + l0] + q[MQ_LAM * st + lm]);
            double h[NS];
            tr::species_enthalpy(T, h);
            double tmp = 0;
#pragma unroll
            for (int n = 0; n < NS; ++n) tmp += J[n] * h[n];
            const double dTn = (Tr - Tl) * dxn;
            double dn[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) dn[a] = (q[(MQ_U + a) * st + l0] - q[(MQ_U + a) * st + lm]) * dxn;
            double gd[3] = {0, 0, 0}, gt[3] = {0, 0, 0};
#pragma unroll
            for (int T2 = 0; T2 < 3; ++T2) {
                if (T2 == D) continue;
                const long ts = T2 == 0 ? 1 : (T2 == 1 ? (long)B.nxg : (long)B.nxg * B.nyg);
                const double h4 = 0.25 * B.dxinv[T2];
                const double* ud = q + (MQ_U + D) * st;
                const double* ut = q + (MQ_U + T2) * st;
                gd[T2] = (ud[l0 + ts] + ud[lm + ts] - ud[l0 - ts] - ud[lm - ts]) * h4;
                gt[T2] = (ut[l0 + ts] + ut[lm + ts] - ut[l0 - ts] - ut[lm - ts]) * h4;
            }
            const double divu = (D == 0 ? dn[0] : gt[0]) + (D == 1 ? dn[1] : gt[1]) + (D == 2 ? dn[2] : gt[2]);
            double tau[3];
#pragma unroll
            for (int a = 0; a < 3; ++a)
                tau[a] = a == D ? muf * (2.0 * dn[D] - m.a.ph.twoThirds * divu) : muf * (gd[a] + dn[a]);
            F[UMX] += -tau[0];
            F[UMY] += -tau[1];
            F[UMZ] += -tau[2];
            F[UEDEN] += -0.5 * ((q[MQ_U * st + l0] + q[MQ_U * st + lm]) * tau[0] +
                                (q[MQ_V * st + l0] + q[MQ_V * st + lm]) * tau[1] +
                                (q[MQ_W * st + l0] + q[MQ_W * st + lm]) * tau[2]) - lamf * dTn - tmp;
        }
        double* Fo = m.mF[D] + B.off_s / NC * NF + l0;
#pragma unroll
        for (int n = 0; n < NF; ++n) Fo[n * st] = F[n];
    }
}

// divop of the 14 face fluxes (Kernels.H:84-97) and the stage combination of
// state and species (MultiFab::LinComb / Saxpy of FlowAdvance.cpp:59-86)
__global__ __launch_bounds__(256) void k_mdiv(const BoxDesc* __restrict__ boxes, const int4* __restrict__ tiles,
                                              MultiArgs m)
{
    if (ctl_done(m.a.ctl)) return;
    const int4 t = tiles[blockIdx.x];
    if ((int)threadIdx.x >= t.z) return;
    const BoxDesc B = boxes[t.x];
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1;
    const int cc = t.y + threadIdx.x;
    const int i = B.lo[0] + cc % nx, j = B.lo[1] + (cc / nx) % ny, k = B.lo[2] + cc / (nx * ny);
    const long st = B.nptg, l0 = gl(B, i, j, k);
    const long sy = B.nxg, sz = (long)B.nxg * B.nyg;
    const long f0 = B.off_s / NC * NF + l0;
    const double* Fx = m.mF[0] + f0;
    const double* Fy = m.mF[1] + f0;
    const double* Fz = m.mF[2] + f0;
    double L[NF];
#pragma unroll
    for (int n = 0; n < NF; ++n)
        L[n] = B.dxinv[0] * (Fx[n * st] - Fx[n * st + 1]) + B.dxinv[1] * (Fy[n * st] - Fy[n * st + sy]) +
               B.dxinv[2] * (Fz[n * st] - Fz[n * st + sz]);
    StageArgs a = m.a;
    a.flags &= ~OCFD_GRAVITY;
    finish_cell(a, B, i, j, k, L);
    const long p0 = B.off_s / NC * NS + l0;
    if (m.pnew) {
        const double cdt = a.ctl ? a.c_dt * a.ctl->dt : a.c_dt;
#pragma unroll
        for (int n = 0; n < NS; ++n) {
            const double v = m.pold ? a.c_cur * m.pcur[p0 + n * st] + a.c_old * m.pold[p0 + n * st]
                                    : a.c_cur * m.pcur[p0 + n * st];
            m.pnew[p0 + n * st] = v + cdt * L[NC + n];
        }
    } else {
        const long d0 = B.off_d / NC * NS + didx(B, i, j, k) - B.off_d;
#pragma unroll
        for (int n = 0; n < NS; ++n) m.pacc[d0 + n * B.nptv] = L[NC + n];
    }
}

// ghost fill of a species MultiFab: every component takes the scalar BC of
// the density (Source/Setup.cpp:241-247); ext_dir sides get the level's
// inflow species when it has them, else they are left alone
__global__ __launch_bounds__(256) void k_fill_spec(const BoxDesc* __restrict__ boxes, const GeomDesc* __restrict__ geoms,
                                                   const int4* __restrict__ tiles, double* __restrict__ P,
                                                   const ocfd_step_ctl* __restrict__ ctl)
{
    if (ctl_done(ctl)) return;
    const int4 t = tiles[blockIdx.x];
    if ((int)threadIdx.x >= t.z) return;
    const BoxDesc B = boxes[t.x];
    const GeomDesc& G = geoms[B.geom];
    int i, j, k;
    ghost_cell(B, t.y + threadIdx.x, t.w, i, j, k);
    const long dst = B.off_s / NC * NS + gl(B, i, j, k);
    int iv[3] = {i, j, k}, outside[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (G.periodic[d]) {
            const int len = G.dom_hi[d] - G.dom_lo[d] + 1;
            int r = (iv[d] - G.dom_lo[d]) % len;
            if (r < 0) r += len;
            iv[d] = G.dom_lo[d] + r;
        }
        outside[d] = iv[d] < G.dom_lo[d] ? -1 : (iv[d] > G.dom_hi[d] ? 1 : 0);
    }
    int src[3] = {iv[0], iv[1], iv[2]};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        if (!outside[d]) continue;
        const int bc = G.bc[0][d][outside[d] > 0];
        const int edge = outside[d] < 0 ? G.dom_lo[d] : G.dom_hi[d];
        if (bc == BC_FOEXT) src[d] = edge;
        else if (bc == BC_EVEN || bc == BC_ODD) src[d] = outside[d] < 0 ? 2 * edge - src[d] - 1 : 2 * edge - src[d] + 1;
        else if (bc == BC_EXT) {
            if (G.has_spec_in) {
#pragma unroll
                for (int n = 0; n < NS; ++n) P[dst + n * B.nptg] = G.spec_in[n];
            }
            return;
        }
    }
    if (src[0] < G.dom_lo[0] || src[0] > G.dom_hi[0] || src[1] < G.dom_lo[1] || src[1] > G.dom_hi[1] ||
        src[2] < G.dom_lo[2] || src[2] > G.dom_hi[2])
        return;
    const int o = level_owner(G, boxes, src[0], src[1], src[2]);
    if (o < 0) return;
    const BoxDesc& O = boxes[o];
    const long s0 = O.off_s / NC * NS + gl(O, src[0], src[1], src[2]);
#pragma unroll
    for (int n = 0; n < NS; ++n) P[dst + n * B.nptg] = P[s0 + n * O.nptg];
}


// ---------------------------------------------------------------------------
// EB-CHEM boxes: EBR::ebchem_compute_dSdt_box (Source/CHEM/FlowAdvanceBox.cpp:16-525)
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool mconn(uint32_t f, int i, int j, int k)
{
    return (f >> (5 + (i + 1) + 3 * (j + 1) + 9 * (k + 1))) & 1u;
}

__device__ __forceinline__ double mminmod(double a, double b)
{
    if (a * b > 0) return fabs(a) > fabs(b) ? b : a;
    return 0.0;
}

// eb_recon (Source/EB/EBkernels.H:10-71) of one component x (stride ds) at the
// low face of cell l0 with the right cell's flag f
__device__ __forceinline__ void mrecon(const double* x, long l0, long ds, bool cm, bool cm2, bool cp, double& ql,
                                       double& qr)
{
    const double q0 = x[l0], qm = x[l0 - ds];
    ql = qm;
    qr = q0;
    if (cm) {
        if (cm2) ql = qm + 0.5 * mminmod(q0 - qm, qm - x[l0 - 2 * ds]);
        if (cp) qr = q0 - 0.5 * mminmod(x[l0 + ds] - q0, q0 - qm);
    }
}

// The mass flux of Van Leer's splitting (B. van Leer, Lecture Notes in
// Physics 170 (1982) 507-512) for a partial density r moving with normal
// velocity U at sound speed c: sgn = +1 the forward part, -1 the backward
// part (the whole flux r U when supersonic towards the side, the split
// r c sgn (1 + sgn M)^2 / 4 when subsonic, else nothing)
__device__ __forceinline__ double vl_mass_part(double r, double U, double c, double sgn)
{
    const double M = U / c;
    if (sgn * M >= 1.0) return r * U;
    if (fabs(M) < 1.0) {
        const double w = 1.0 + sgn * M;
        return r * c * (sgn * 0.250 * w * w);
    }
    return 0.0;
}

// species face flux of a cut box: MUSCL states of u_D, c and rho_k, the Van
// Leer mass-flux split of each partial density (the species flux of
// FlowAdvanceBox.cpp:174-223), Fickian flux when viscous
template <int D>
__global__ __launch_bounds__(256) void k_ebchem_face(const BoxDesc* __restrict__ boxes, const int4* __restrict__ tiles,
                                                    MultiArgs m, const uint32_t* __restrict__ ebflag)
{
    if (ctl_done(m.a.ctl)) return;
    const int4 t = tiles[blockIdx.x];
    for (int c = threadIdx.x; c < t.z; c += 256) {   // tiles hold at most 256 cells
        const BoxDesc B = boxes[t.x];
        const int di = D == 0, dj = D == 1, dk = D == 2;
        const int nx = B.hi[0] - B.lo[0] + 1 + di, ny = B.hi[1] - B.lo[1] + 1 + dj;
        const int cc = t.y + c;
        const int i = B.lo[0] + cc % nx, j = B.lo[1] + (cc / nx) % ny, k = B.lo[2] + cc / (nx * ny);
        const long st = B.nptg;
        const long ds = D == 0 ? 1 : (D == 1 ? (long)B.nxg : (long)B.nxg * B.nyg);
        const long l0 = gl(B, i, j, k);
        const double* q = m.mq + B.off_s / NC * MQ_N;
        const double* P = m.pcur + B.off_s / NC * NS;
        const uint32_t f = ebflag[B.off_e + l0];
        const bool cm = mconn(f, -di, -dj, -dk), cm2 = mconn(f, -2 * di, -2 * dj, -2 * dk), cp = mconn(f, di, dj, dk);
        double ul, ur, cL, cR;
        mrecon(q + (MQ_U + D) * st, l0, ds, cm, cm2, cp, ul, ur);
        mrecon(q + MQ_C * st, l0, ds, cm, cm2, cp, cL, cR);
        double F[NS];
#pragma unroll
        for (int n = 0; n < NS; ++n) {
            double rl, rr;
            mrecon(P + n * st, l0, ds, cm, cm2, cp, rl, rr);
            F[n] = vl_mass_part(rl, ul, cL, 1.0) + vl_mass_part(rr, ur, cR, -1.0);
        }
        if (m.a.flags & OCFD_VISC) {
            const long lm = l0 - ds;
            const double dxn = B.dxinv[D];
            const double rhor = q[MQ_R * st + l0], rhol = q[MQ_R * st + lm];
            const double rho = 0.5 * (rhol + rhor);
#pragma unroll
            for (int n = 0; n < NS; ++n) {
                const double Ytr = P[n * st + l0] / rhor;
                const double Ytl = P[n * st + lm] / rhol;
                const double dn = 0.5 * (q[(MQ_D + n) * st + l0] + q[(MQ_D + n) * st + lm]);
                F[n] += rho * dn * (Ytl - Ytr) * dxn;
            }
        }
        double* Fo = m.mF[D] + B.off_s / NC * NF + l0 + NC * st;
#pragma unroll
        for (int n = 0; n < NS; ++n) Fo[n * st] = F[n];
    }
}

// per valid cell of cut and covered boxes: the flow dSdt is 0 (covered boxes,
// Source/CHEM/FlowAdvance.cpp:127-129; cut boxes, FlowAdvanceBox.cpp:501-524);
// the species divergence with area fractions and the wall term (:414-460)
__global__ __launch_bounds__(256) void k_ebchem_div(const BoxDesc* __restrict__ boxes, const int4* __restrict__ tiles,
                                                    MultiArgs m, const uint32_t* __restrict__ ebflag,
                                                    const double* __restrict__ ebgeo)
{
    if (ctl_done(m.a.ctl)) return;
    const int4 t = tiles[blockIdx.x];
    if ((int)threadIdx.x >= t.z) return;
    const BoxDesc B = boxes[t.x];
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1;
    const int cc = t.y + threadIdx.x;
    const int i = B.lo[0] + cc % nx, j = B.lo[1] + (cc / nx) % ny, k = B.lo[2] + cc / (nx * ny);
    const long st = B.nptg, l0 = gl(B, i, j, k);
    double L[NF];
#pragma unroll
    for (int n = 0; n < NF; ++n) L[n] = 0.0;
    if (B.type == 1) {
        const long sy = B.nxg, sz = (long)B.nxg * B.nyg;
        const uint32_t f = ebflag[B.off_e + l0];
        const double* geo = ebgeo + B.off_e * 4;
        const double vf = geo[l0];
        const double* ax = geo + st;
        const double* ay = geo + 2 * st;
        const double* az = geo + 3 * st;
        const long f0 = B.off_s / NC * NF + l0 + NC * st;
        const double* Fx = m.mF[0] + f0;
        const double* Fy = m.mF[1] + f0;
        const double* Fz = m.mF[2] + f0;
        const double tmp = B.dxinv[0] / vf;
        double Ls[NS];
        if ((f & 3u) == 3u) {
#pragma unroll
            for (int n = 0; n < NS; ++n) Ls[n] = 0.0;
        } else if ((f & 3u) == 0u) {
#pragma unroll
            for (int n = 0; n < NS; ++n)
                Ls[n] = B.dxinv[0] * (Fx[n * st + 1] - Fx[n * st]) + B.dxinv[1] * (Fy[n * st + sy] - Fy[n * st]) +
                        B.dxinv[2] * (Fz[n * st + sz] - Fz[n * st]);
        } else {
            const double axm = ax[l0], axp = ax[l0 + 1], aym = ay[l0], ayp = ay[l0 + sy], azm = az[l0], azp = az[l0 + sz];
#pragma unroll
            for (int n = 0; n < NS; ++n)
                Ls[n] = tmp * (axp * Fx[n * st + 1] - axm * Fx[n * st] + ayp * Fy[n * st + sy] - aym * Fy[n * st] +
                               azp * Fz[n * st + sz] - azm * Fz[n * st]);
            if (m.a.flags & OCFD_VISC) {
                const double* q = m.mq + B.off_s / NC * MQ_N + l0;
                const double* P = m.pcur + B.off_s / NC * NS + l0;
                const double apnorm =
                    sqrt((axm - axp) * (axm - axp) + (aym - ayp) * (aym - ayp) + (azm - azp) * (azm - azp));
                const double un = q[MQ_U * st] * (axm - axp) + q[MQ_V * st] * (aym - ayp) + q[MQ_W * st] * (azm - azp);
                const double lambda0 = fabs(un) + q[MQ_C * st] * apnorm;
#pragma unroll
                for (int n = 0; n < NS; ++n) {
                    const double fp = 0.5 * (P[n * st] * un + lambda0 * P[n * st]);
                    const double fm = -0.5 * (P[n * st] * un + lambda0 * P[n * st]);
                    Ls[n] += (fp + fm) * tmp;
                }
            }
        }
#pragma unroll
        for (int n = 0; n < NS; ++n) L[NC + n] = -1.0 * Ls[n];
    }
    StageArgs a = m.a;
    a.flags &= ~OCFD_GRAVITY;
    finish_cell(a, B, i, j, k, L);
    const long p0 = B.off_s / NC * NS + l0;
    if (m.pnew) {
        const double cdt = a.ctl ? a.c_dt * a.ctl->dt : a.c_dt;
#pragma unroll
        for (int n = 0; n < NS; ++n) {
            const double v = m.pold ? a.c_cur * m.pcur[p0 + n * st] + a.c_old * m.pold[p0 + n * st]
                                    : a.c_cur * m.pcur[p0 + n * st];
            m.pnew[p0 + n * st] = v + cdt * L[NC + n];
        }
    } else {
        const long d0 = B.off_d / NC * NS + didx(B, i, j, k) - B.off_d;
#pragma unroll
        for (int n = 0; n < NS; ++n) m.pacc[d0 + n * B.nptv] = L[NC + n];
    }
}

void launch_multi(const LevelDev& L, const MultiArgs& m, hipStream_t s)
{
    if (L.ntile_p) hipLaunchKernelGGL(k_mprims, dim3(L.ntile_p), dim3(256), 0, s, L.d_boxes, L.d_tile_p, m);
    if (L.ntile_mf[0]) {
        hipLaunchKernelGGL(k_mface<0>, dim3(L.ntile_mf[0]), dim3(256), 0, s, L.d_boxes, L.d_tile_mf[0], m);
        hipLaunchKernelGGL(k_mface<1>, dim3(L.ntile_mf[1]), dim3(256), 0, s, L.d_boxes, L.d_tile_mf[1], m);
        hipLaunchKernelGGL(k_mface<2>, dim3(L.ntile_mf[2]), dim3(256), 0, s, L.d_boxes, L.d_tile_mf[2], m);
    }
    if (L.ntile_md) hipLaunchKernelGGL(k_mdiv, dim3(L.ntile_md), dim3(256), 0, s, L.d_boxes, L.d_tile_md, m);
    // cut and covered boxes
    if (L.ntile_eq) {
        hipLaunchKernelGGL(k_mprims, dim3(L.ntile_eq), dim3(256), 0, s, L.d_boxes, L.d_tile_eq, m);
        hipLaunchKernelGGL(k_ebchem_face<0>, dim3(L.ntile_ef[0]), dim3(256), 0, s, L.d_boxes, L.d_tile_ef[0], m,
                           (const uint32_t*)L.d_ebflag);
        hipLaunchKernelGGL(k_ebchem_face<1>, dim3(L.ntile_ef[1]), dim3(256), 0, s, L.d_boxes, L.d_tile_ef[1], m,
                           (const uint32_t*)L.d_ebflag);
        hipLaunchKernelGGL(k_ebchem_face<2>, dim3(L.ntile_ef[2]), dim3(256), 0, s, L.d_boxes, L.d_tile_ef[2], m,
                           (const uint32_t*)L.d_ebflag);
    }
    if (L.ntile_ed)
        hipLaunchKernelGGL(k_ebchem_div, dim3(L.ntile_ed), dim3(256), 0, s, L.d_boxes, L.d_tile_ed, m,
                           (const uint32_t*)L.d_ebflag, (const double*)L.d_ebgeo);
}

void launch_fill_spec(const LevelDev& L, double* P, const ocfd_step_ctl* ctl, hipStream_t s)
{
    if (L.ntile_g)
        hipLaunchKernelGGL(k_fill_spec, dim3(L.ntile_g), dim3(256), 0, s, L.d_boxes, L.d_geom, L.d_tile_g, P, ctl);
}

// every cell of every box (valid and ghost) set to one state and species
// set: the premixed initial data and inflow of the reacting channel
struct UniformState {
    double u[NC];
    double r[NS];
};

// grid: x over the grown cells of a box, y over the boxes
__global__ __launch_bounds__(256) void k_init_uniform(const BoxDesc* __restrict__ boxes, double* __restrict__ S,
                                                      double* __restrict__ P, UniformState v)
{
    const BoxDesc B = boxes[blockIdx.y];
    const long p0 = B.off_s / NC * NS;
    for (long c = (long)blockIdx.x * 256 + threadIdx.x; c < B.nptg; c += (long)gridDim.x * 256) {
#pragma unroll
        for (int n = 0; n < NC; ++n) S[B.off_s + n * B.nptg + c] = v.u[n];
#pragma unroll
        for (int n = 0; n < NS; ++n) P[p0 + n * B.nptg + c] = v.r[n];
    }
}

// A <- B for the forward-Euler flow step under device control
__global__ __launch_bounds__(256) void k_copy_ctl_m(double* __restrict__ dst, const double* __restrict__ src, long n,
                                                    const ocfd_step_ctl* __restrict__ ctl)
{
    if (ctl_done(ctl)) return;
    for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)gridDim.x * 256) dst[i] = src[i];
}

// Exec/5_CHEMSOD/prob.H:11-82 with prob_parm.H defaults: H2:O2 = 0.667:0.333
// at rest; 20 atm and 1500 K for x < 0.0004, 0.2 atm and 295 K elsewhere
__global__ __launch_bounds__(256) void k_init_chemsod(const BoxDesc* __restrict__ boxes, const int4* __restrict__ tiles,
                                                      double* __restrict__ S, double* __restrict__ P, double prob_lo,
                                                      double dx)
{
#pragma clang fp contract(off)
    const int4 t = tiles[blockIdx.x];
    if ((int)threadIdx.x >= t.z) return;
    const BoxDesc B = boxes[t.x];
    const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1;
    const int cc = t.y + threadIdx.x;
    const int i = B.lo[0] + cc % nx, j = B.lo[1] + (cc / nx) % ny, k = B.lo[2] + cc / (nx * ny);
    const long st = B.nptg, l0 = gl(B, i, j, k);
    double Pt = 0.2 * 101325.0, temp = 295;
    double X[NS] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, Y[NS], rhoi[NS];
    X[0] = 0.667;
    X[1] = 0.333;
    const double x = prob_lo + (i + 0.5) * dx;
    if (x < 0.0004) { Pt = 20 * 101325.0; temp = 1500; }
    // CKXTY, CKRHOY (LiDryer.H:21-66)
    double XW = 0.0;
    for (int n = 0; n < NS; ++n) XW += X[n] * chem::mw[n];
    for (int n = 0; n < NS; ++n) Y[n] = X[n] * chem::mw[n] / XW;
    double tmp[NS], YOW = 0;
    for (int n = 0; n < NS; ++n) tmp[n] = Y[n] / chem::mw[n];
    for (int n = 0; n < NS; ++n) YOW += tmp[n];
    const double rho = Pt / (chem::Ru * temp * YOW);
    for (int n = 0; n < NS; ++n) rhoi[n] = rho * Y[n];
    const double et = chem::ckubms(temp, rhoi);
    double* s = S + B.off_s + l0;
    s[0] = rho; s[st] = 0.0; s[2 * st] = 0.0; s[3 * st] = 0.0; s[4 * st] = et;
    double* p = P + B.off_s / NC * NS + l0;
    for (int n = 0; n < NS; ++n) p[n * st] = rho * Y[n];
}

// Scratch of the allocation-free multispecies entry point: the 19-component
// primitive/transport buffer and the three 14-component face-flux buffers of
// one level, laid out as ensure_multi lays out its allocations, in a
// file-scope device array.  Each GPU holds its own copy from the first use of
// this module on it until the process ends.  Host code never uses the array's
// name as an address: the per-device address comes from hipGetSymbolAddress.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
__device__ __align__(256) double scratch[kNoAllocScratchBytes / sizeof(double)];

// bytes of scratch the multispecies path needs on level L
inline size_t multi_scratch_bytes(const LevelDev& L)
{
    return sizeof(double) * (size_t)(MQ_N + 3 * NF) * (size_t)(L.mf_state / NC);
}

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and kept.
static int noalloc_scratch(double** out)
{
    static std::mutex mu;
    static std::map<int, double*> addr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error("ocfd_compute_dsdt_multi_noalloc: no current device");
        return OCFD_EHIP;
    }
    std::lock_guard<std::mutex> lk(mu);
    auto it = addr.find(dev);
    if (it == addr.end()) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("ocfd_compute_dsdt_multi_noalloc: hipGetSymbolAddress: ") + hipGetErrorString(e));
            return OCFD_EHIP;
        }
        it = addr.emplace(dev, (double*)p).first;
    }
    *out = it->second;
    return OCFD_OK;
}

} // namespace ocfd

using namespace ocfd;

extern "C" {

static int mcheck(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        return OCFD_EHIP;
    }
    return OCFD_OK;
}

static int ensure_multi(LevelDev& L)
{
    if (L.d_mq) return OCFD_OK;
    const long ng = L.mf_state / NC;
    double *mq = nullptr, *mF[3] = {nullptr, nullptr, nullptr};
    bool ok = dev_alloc((void**)&mq, sizeof(double) * MQ_N * ng);
    for (int d = 0; d < 3 && ok; ++d) ok = dev_alloc((void**)&mF[d], sizeof(double) * NF * ng);
    if (!ok) {   // all or nothing: a later call tries again from scratch
        dev_free({mq, mF[0], mF[1], mF[2]});
        set_error("multispecies flow: out of device memory");
        return OCFD_ENOMEM;
    }
    L.d_mq = mq;
    for (int d = 0; d < 3; ++d) L.d_mF[d] = mF[d];
    L.dev_bytes += (long)sizeof(double) * (MQ_N + 3 * NF) * ng;
    return OCFD_OK;
}

static MultiArgs margs(LevelDev& L, const double* S, const double* P, const double* So, const double* Po, double* Sn,
                       double* Pn, double* acc, double* pacc, double cc, double co, double cdt, const ocfd_parm* parm,
                       int flags)
{
    MultiArgs m{};
    m.a.scur = S; m.a.sold = So; m.a.snew = Sn; m.a.acc = acc;
    m.a.c_cur = cc; m.a.c_old = co; m.a.c_dt = cdt;
    m.a.flags = flags & OCFD_VISC;
    m.a.ph = phys_of(*parm);
    m.a.ctl = nullptr;
    m.pcur = P; m.pold = Po; m.pnew = Pn; m.pacc = pacc;
    m.mq = L.d_mq;
    for (int d = 0; d < 3; ++d) m.mF[d] = L.d_mF[d];
    return m;
}

int ocfd_fill_boundary_spec(ocfd_level lev, double* Spec, double time, void* stream)
{
    if (lev && lev->dev.fill_gz != NG) {
        set_error("multispecies stepping needs the full ghost fill depth (ocfd_level_set_planar off)");
        return OCFD_EINVAL;
    }
    (void)time;
    if (!lev || !Spec) { set_error("ocfd_fill_boundary_spec: invalid arguments"); return OCFD_EINVAL; }
    launch_fill_spec(lev->dev, Spec, nullptr, (hipStream_t)stream);
    return mcheck("ocfd_fill_boundary_spec");
}

int ocfd_compute_dsdt_multi(ocfd_level lev, const double* S, const double* Spec, double* dSdt, double* dSdt_spec,
                            double dt, const ocfd_parm* parm, int flags, void* stream)
{
    if (lev && lev->dev.fill_gz != NG) {
        set_error("multispecies stepping needs the full ghost fill depth (ocfd_level_set_planar off)");
        return OCFD_EINVAL;
    }
    (void)dt;
    if (!lev || !S || !Spec || !dSdt || !dSdt_spec || !parm) {
        set_error("ocfd_compute_dsdt_multi: invalid arguments");
        return OCFD_EINVAL;
    }
    int rc = ensure_multi(lev->dev);
    if (rc) return rc;
    MultiArgs m = margs(lev->dev, S, Spec, nullptr, nullptr, nullptr, nullptr, dSdt, dSdt_spec, 1.0, 0.0, 0.0, parm,
                        flags);
    launch_multi(lev->dev, m, (hipStream_t)stream);
    return mcheck("ocfd_compute_dsdt_multi");
}

int ocfd_compute_dsdt_multi_noalloc(ocfd_level lev, const double* S, const double* Spec, double* dSdt,
                                    double* dSdt_spec, double dt, const ocfd_parm* parm, int flags, void* stream)
{
    if (lev && lev->dev.fill_gz != NG) {
        set_error("multispecies stepping needs the full ghost fill depth (ocfd_level_set_planar off)");
        return OCFD_EINVAL;
    }
    (void)dt;
    if (!lev || !S || !Spec || !dSdt || !dSdt_spec || !parm) {
        set_error("ocfd_compute_dsdt_multi_noalloc: invalid arguments");
        return OCFD_EINVAL;
    }
    const LevelDev& L = lev->dev;
    const size_t need = multi_scratch_bytes(L);
    if (need > kNoAllocScratchBytes) {
        set_error("ocfd_compute_dsdt_multi_noalloc: the level needs " + std::to_string(need) +
                  " bytes of scratch; the allocation-free entry point holds " + std::to_string(kNoAllocScratchBytes) +
                  " (use ocfd_compute_dsdt_multi)");
        return OCFD_EINVAL;
    }
    double* base = nullptr;
    int rc = noalloc_scratch(&base);
    if (rc) return rc;
    MultiArgs m = margs(lev->dev, S, Spec, nullptr, nullptr, nullptr, nullptr, dSdt, dSdt_spec, 1.0, 0.0, 0.0, parm,
                        flags);
    // the layout of ensure_multi: primitives, then the x, y, z face fluxes;
    // every element is written before it is read, so nothing is cleared
    const long ng = L.mf_state / NC;
    m.mq = base;
    for (int d = 0; d < 3; ++d) m.mF[d] = base + (long)MQ_N * ng + (long)d * NF * ng;
    launch_multi(L, m, (hipStream_t)stream);
    return mcheck("ocfd_compute_dsdt_multi_noalloc");
}

int ocfd_rk_stage_multi(ocfd_level lev, const double* S_cur, const double* Spec_cur, const double* S_old,
                        const double* Spec_old, double* S_new, double* Spec_new, double c_cur, double c_old,
                        double c_dt, const ocfd_parm* parm, int flags, void* stream)
{
    if (lev && lev->dev.fill_gz != NG) {
        set_error("multispecies stepping needs the full ghost fill depth (ocfd_level_set_planar off)");
        return OCFD_EINVAL;
    }
    if (!lev || !S_cur || !Spec_cur || !S_new || !Spec_new || !parm || S_cur == S_new || Spec_cur == Spec_new ||
        (!S_old) != (!Spec_old)) {
        set_error("ocfd_rk_stage_multi: invalid arguments");
        return OCFD_EINVAL;
    }
    int rc = ensure_multi(lev->dev);
    if (rc) return rc;
    MultiArgs m = margs(lev->dev, S_cur, Spec_cur, S_old, Spec_old, S_new, Spec_new, nullptr, nullptr, c_cur, c_old,
                        c_dt, parm, flags);
    launch_multi(lev->dev, m, (hipStream_t)stream);
    return mcheck("ocfd_rk_stage_multi");
}

int ocfd_advance_multi(ocfd_level lev, double* A, double* Aspec, double* B, double* Bspec, double time, double dt,
                       int scheme, const ocfd_parm* parm, int flags, void* stream)
{
    if (lev && lev->dev.fill_gz != NG) {
        set_error("multispecies stepping needs the full ghost fill depth (ocfd_level_set_planar off)");
        return OCFD_EINVAL;
    }
    if (!lev || !A || !Aspec || !B || !Bspec || !parm || scheme < 0 || scheme > 1) {
        set_error("ocfd_advance_multi: invalid arguments (scheme 0 Euler or 1 RK2)");
        return OCFD_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    LevelDev& L = lev->dev;
    int rc;
    if ((rc = ocfd_fill_boundary(lev, A, time, s)) || (rc = ocfd_fill_boundary_spec(lev, Aspec, time, s))) return rc;
    if ((rc = ocfd_rk_stage_multi(lev, A, Aspec, nullptr, nullptr, B, Bspec, 1.0, 0.0, dt, parm, flags, s))) return rc;
    if (scheme == 0) {
        if (hipMemcpyAsync(A, B, sizeof(double) * L.mf_state, hipMemcpyDeviceToDevice, s) != hipSuccess ||
            hipMemcpyAsync(Aspec, Bspec, sizeof(double) * (L.mf_state / NC * NS), hipMemcpyDeviceToDevice, s) !=
                hipSuccess)
            return mcheck("ocfd_advance_multi");
        return OCFD_OK;
    }
    if ((rc = ocfd_fill_boundary(lev, B, time + dt, s)) || (rc = ocfd_fill_boundary_spec(lev, Bspec, time + dt, s)))
        return rc;
    return ocfd_rk_stage_multi(lev, B, Bspec, A, Aspec, A, Aspec, 0.5, 0.5, 0.5 * dt, parm, flags, s);
}

} // extern "C"

namespace {

constexpr double kChemEstdt = 1e-8;   // the CHEM build's ebr_estdt (Source/Advance.cpp:54-57)

struct ChemStep {
    ocfd_level lev;
    double *A, *Aspec, *B, *Bspec;
    ocfd_step_ctl* ctl;
    int scheme, level, flags, renorm, nsteps;
    const ocfd_parm* parm;
};

// One Strang-split EBR::advance of a multispecies level with dt and time from
// ctl (Source/Advance.cpp:152-177 with flow_advance_multi,
// Source/CHEM/FlowAdvance.cpp:9-90), the renormalisation of post_timestep and
// the dt update -- the launches of ocfd_advance_chem, under device control
int chem_step(const ChemStep& c, hipStream_t s)
{
    LevelDev& L = c.lev->dev;
    const int iter = c.level <= 2 ? 4 / (1 << c.level) : 1;
    const double smallr = c.parm->smallr;
    for (int n = 0; n < iter; ++n) launch_chem_level(L, c.A, c.Aspec, 0.0, smallr, c.ctl, iter, s);
    auto stage = [&](const double* S, const double* P, const double* So, const double* Po, double* Sn, double* Pn,
                     double cc, double co, double cdt) {
        MultiArgs m = margs(L, S, P, So, Po, Sn, Pn, nullptr, nullptr, cc, co, cdt, c.parm, c.flags);
        m.a.ctl = c.ctl;
        launch_multi(L, m, s);
    };
    launch_fill_ctl(L, c.A, c.ctl, 0.0, s);
    launch_fill_spec(L, c.Aspec, c.ctl, s);
    stage(c.A, c.Aspec, nullptr, nullptr, c.B, c.Bspec, 1.0, 0.0, 1.0);
    if (c.scheme == 0) {
        const long n = L.mf_state, np = L.mf_state / NC * NS;
        hipLaunchKernelGGL(k_copy_ctl_m, dim3((unsigned)std::min<long>((n + 255) / 256, 4096)), dim3(256), 0, s, c.A,
                           c.B, n, (const ocfd_step_ctl*)c.ctl);
        hipLaunchKernelGGL(k_copy_ctl_m, dim3((unsigned)std::min<long>((np + 255) / 256, 4096)), dim3(256), 0, s,
                           c.Aspec, c.Bspec, np, (const ocfd_step_ctl*)c.ctl);
    } else {
        launch_fill_ctl(L, c.B, c.ctl, 1.0, s);
        launch_fill_spec(L, c.Bspec, c.ctl, s);
        stage(c.B, c.Bspec, c.A, c.Aspec, c.A, c.Aspec, 0.5, 0.5, 0.5);
    }
    for (int n = 0; n < iter; ++n) launch_chem_level(L, c.A, c.Aspec, 0.0, smallr, c.ctl, iter, s);
    if (c.renorm) launch_renorm(L, c.A, c.Aspec, c.ctl, s);
    launch_step_end(L, c.A, phys_of(*c.parm), c.ctl, kChemEstdt, s);
    return mcheck("ocfd_step_device_chem");
}

int chem_steps_body(void* p, hipStream_t s)
{
    const ChemStep& c = *(const ChemStep*)p;
    int rc = OCFD_OK;
    for (int n = 0; n < c.nsteps && rc == OCFD_OK; ++n) rc = chem_step(c, s);
    return rc;
}

bool chem_step_ok(const ChemStep& c, const char* what)
{
    if (c.lev && c.lev->dev.fill_gz != NG) {
        set_error("multispecies stepping needs the full ghost fill depth (ocfd_level_set_planar off)");
        return false;
    }
    const bool ok = c.lev && c.A && c.Aspec && c.B && c.Bspec && c.ctl && c.parm && c.level >= 0 &&
                    (c.scheme == 0 || c.scheme == 1) && c.A != c.B && c.Aspec != c.Bspec;
    if (!ok) set_error(std::string(what) + ": invalid arguments (scheme 0 Euler or 1 RK2)");
    return ok;
}

} // namespace

extern "C" {

int ocfd_ctl_start_chem(ocfd_level lev, const double* A, ocfd_step_ctl* ctl, double time, double cfl,
                        double stop_time, int max_steps, const ocfd_parm* parm, void* stream)
{
    if (!lev || !A || !ctl || !parm || !(cfl > 0.0)) {
        set_error("ocfd_ctl_start_chem: invalid arguments");
        return OCFD_EINVAL;
    }
    launch_ctl_start(lev->dev, A, phys_of(*parm), ctl, time, cfl, stop_time, max_steps, kChemEstdt,
                     (hipStream_t)stream);
    return mcheck("ocfd_ctl_start_chem");
}

int ocfd_step_device_chem(ocfd_level lev, double* A, double* Aspec, double* B, double* Bspec, ocfd_step_ctl* ctl,
                          int scheme, int level, const ocfd_parm* parm, int flags, int renorm, void* stream)
{
    ChemStep c{lev, A, Aspec, B, Bspec, ctl, scheme, level, flags, renorm, 1, parm};
    if (!chem_step_ok(c, "ocfd_step_device_chem")) return OCFD_EINVAL;
    int rc = ensure_multi(lev->dev);
    if (rc) return rc;
    return chem_step(c, (hipStream_t)stream);
}

int ocfd_graph_capture_steps_chem(ocfd_level lev, double* A, double* Aspec, double* B, double* Bspec,
                                  ocfd_step_ctl* ctl, int scheme, int level, const ocfd_parm* parm, int flags,
                                  int renorm, int nsteps, void* stream, ocfd_graph* out)
{
    ChemStep c{lev, A, Aspec, B, Bspec, ctl, scheme, level, flags, renorm, nsteps, parm};
    if (!chem_step_ok(c, "ocfd_graph_capture_steps_chem")) return OCFD_EINVAL;
    if (nsteps <= 0 || !out || !stream) {
        set_error("ocfd_graph_capture_steps_chem: invalid arguments (a non-null stream is required)");
        return OCFD_EINVAL;
    }
    int rc = ensure_multi(lev->dev);   // the scratch is allocated before the capture
    if (rc) return rc;
    return capture_graph(lev, (hipStream_t)stream, chem_steps_body, &c, out, "ocfd_graph_capture_steps_chem");
}

int ocfd_advance_chem(ocfd_level lev, double* A, double* Aspec, double* B, double* Bspec, double time, double dt,
                      int level, int scheme, const ocfd_parm* parm, int flags, int renorm, void* stream)
{
    if (lev && lev->dev.fill_gz != NG) {
        set_error("multispecies stepping needs the full ghost fill depth (ocfd_level_set_planar off)");
        return OCFD_EINVAL;
    }
    if (!lev || !A || !Aspec || !B || !Bspec || !parm || level < 0) {
        set_error("ocfd_advance_chem: invalid arguments");
        return OCFD_EINVAL;
    }
    hipStream_t s = (hipStream_t)stream;
    int iter = level <= 2 ? 4 / (1 << level) : 1;
    const double dt1 = 0.5 * dt;
    const double dt2 = dt1 / iter;
    int rc;
    for (int n = 0; n < iter; ++n)
        if ((rc = ocfd_chem_advance(lev, A, Aspec, dt2, parm, s))) return rc;
    if ((rc = ocfd_advance_multi(lev, A, Aspec, B, Bspec, time, dt, scheme, parm, flags, s))) return rc;
    for (int n = 0; n < iter; ++n)
        if ((rc = ocfd_chem_advance(lev, A, Aspec, dt2, parm, s))) return rc;
    if (renorm) return ocfd_species_renorm(lev, A, Aspec, s);
    return OCFD_OK;
}

int ocfd_init_species_uniform(ocfd_level lev, const double* state, const double* spec, double* S, double* Spec)
{
    if (!lev || !state || !spec || !S || !Spec) {
        set_error("ocfd_init_species_uniform: invalid arguments");
        return OCFD_EINVAL;
    }
    const LevelDev& L = lev->dev;
    UniformState v;
    for (int n = 0; n < NC; ++n) v.u[n] = state[n];
    for (int n = 0; n < NS; ++n) v.r[n] = spec[n];
    long maxg = 1;
    for (const BoxDesc& B : L.boxes) maxg = std::max(maxg, B.nptg);
    if (L.nbox > 0) {
        const unsigned gx = (unsigned)std::min<long>((maxg + 255) / 256, 1024);
        hipLaunchKernelGGL(k_init_uniform, dim3(gx, (unsigned)L.nbox), dim3(256), 0, 0, L.d_boxes, S, Spec, v);
    }
    int rc = mcheck("ocfd_init_species_uniform");
    if (rc) return rc;
    if (hipDeviceSynchronize() != hipSuccess) return mcheck("ocfd_init_species_uniform");
    return OCFD_OK;
}

int ocfd_init_species_state(ocfd_level lev, int kind, double* S, double* Spec, const ocfd_parm* parm)
{
    if (!lev || !S || !Spec || !parm || kind != OCFD_CASE_CHEMSOD) {
        set_error("ocfd_init_species_state: invalid arguments");
        return OCFD_EINVAL;
    }
    const LevelDev& L = lev->dev;
    if (hipMemset(S, 0, sizeof(double) * L.mf_state) != hipSuccess ||
        hipMemset(Spec, 0, sizeof(double) * (L.mf_state / NC * NS)) != hipSuccess)
        return mcheck("ocfd_init_species_state");
    if (L.ntile_md || L.ntile_ed) {
        // every box's valid cells: the multispecies and the cut/covered lists
        if (L.ntile_md)
            hipLaunchKernelGGL(k_init_chemsod, dim3(L.ntile_md), dim3(256), 0, 0, L.d_boxes, L.d_tile_md, S, Spec,
                               L.geom.prob_lo[0], L.geom.dx[0]);
        if (L.ntile_ed)
            hipLaunchKernelGGL(k_init_chemsod, dim3(L.ntile_ed), dim3(256), 0, 0, L.d_boxes, L.d_tile_ed, S, Spec,
                               L.geom.prob_lo[0], L.geom.dx[0]);
    }
    int rc = mcheck("ocfd_init_species_state");
    if (rc) return rc;
    if (hipDeviceSynchronize() != hipSuccess) return mcheck("ocfd_init_species_state");
    return OCFD_OK;
}

} // extern "C"
