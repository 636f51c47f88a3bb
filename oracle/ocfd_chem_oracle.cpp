// Li-Dryer H2/air chemistry -- CPU ORACLE (test infrastructure only).
//
// FP64 restatement of the reference's 9-species / 19-reaction kinetics, NASA7
// thermodynamics, temperature inversion, 9x9 Gauss solve, the per-cell
// linearised-implicit chemistry step and the species renormalisation, in the
// reference's operation order (built with -ffp-contract=off).  Pinned by the
// reference's own AMReX-free copy of these functions (Test/chemical_test.cpp,
// compiled into oracle/_ref by oracle/ref.mk), the reference's known answers
// (Exec/7_JET/prob.cpp:18-37), element conservation and a library ODE solve.
//
// Species order (Source/Constants.cpp:13-22): H2 O2 H2O H O OH HO2 H2O2 N2.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>

namespace {

constexpr int NS = 9;
constexpr int NR = 19;
constexpr double Ru = 8.31446261815324;
constexpr double Patm = 101325.0;

// kg/mol, Source/Constants.cpp:13-22
const double mw[NS] = {2.016 * 1e-3, 31.998 * 1e-3, 18.015 * 1e-3, 1.008 * 1e-3, 15.999 * 1e-3,
                       17.007 * 1e-3, 33.006 * 1e-3, 34.014 * 1e-3, 28.014 * 1e-3};

// NASA7 coefficients (low branch T < 1000 K, high branch), Source/Constants.cpp:23-73
const double nasa_lo[NS][7] = {
    {3.29812, 8.24944e-4, -81.43015e-8, -9.47543e-11, 413.48718e-15, -0.01013e+5, -3.29409},
    {3.21294, 11.27486e-4, -57.56150e-8, 131.38770e-11, -876.8554e-15, -0.01005e+5, 6.03474},
    {3.38684, 34.74982e-4, -635.4696e-8, 696.8581e-11, -2506.588e-15, -0.30208e+5, 2.59023},
    {2.5, 0.0, 0.0, 0.0, 0.0, 0.25472e+5, -0.46012},
    {2.94643, -16.3817e-4, 242.10320e-8, -160.284e-11, 389.06961e-15, 0.29148e+5, 2.96399},
    {3.63727, 1.85091e-4, -167.6165e-8, 238.72031e-11, -843.1442e-15, 0.03607e+5, 1.35886},
    {2.97996, 49.96697e-4, -379.0997e-8, 235.41919e-11, -808.9024e-15, 0.00176e+5, 9.22272},
    {3.38875, 65.69226e-4, -14.85013e-8, -462.58060e-11, 2471.515e-15, -0.17663e+5, 6.78536},
    {3.29868, 14.08240e-4, -396.3222e-8, 564.15149e-11, -2444.855e-15, -0.01021e+5, 3.95037}};
const double nasa_hi[NS][7] = {
    {2.99142, 7.00064e-4, -5.63383e-8, -0.92316e-11, 1.58275e-15, -0.00835e+5, -1.35511},
    {3.69758, 6.13520e-4, -12.58842e-8, 1.77528e-11, -1.13644e-15, -0.01234e+5, 3.18917},
    {2.67215, 30.56293e-4, -87.30260e-8, 12.00996e-11, -6.39162e-15, -0.29899e+5, 6.86282},
    {2.5, 0.0, 0.0, 0.0, 0.0, 0.25472e+5, -0.46012},
    {2.54206, -0.27551e-4, -0.31028e-8, 0.45511e-11, -0.43681e-15, 0.29231e+5, 4.92031},
    {2.88273, 10.13974e-4, -22.76877e-8, 2.17468e-11, -0.51263e-15, 0.03887e+5, 5.59571},
    {4.07219, 21.31296e-4, -53.08145e-8, 6.11227e-11, -2.84116e-15, -0.00158e+5, 3.47603},
    {4.57317, 43.36136e-4, -147.4689e-8, 23.48904e-11, -14.31654e-15, -0.18007e+5, 0.50114},
    {2.92664, 14.87977e-4, -56.84761e-8, 10.09704e-11, -6.75335e-15, -0.00923e+5, 5.98053}};

inline double HP(int n) { return nasa_lo[n][5]; }          // Constants.cpp:78
inline double A0(int n) { return nasa_lo[n][0] - 1.0; }    // Constants.cpp:81

// g/(R T), Source/CHEM/LiDryer.H:67-154
void gibbs(double* g, const double* tc)
{
    double T = tc[1];
    double invT = 1 / T;
    for (int s = 0; s < NS; ++s) {
        const double* a = T < 1000 ? nasa_lo[s] : nasa_hi[s];
        g[s] = a[0] * (1.0 - tc[0]) - a[1] / 2.0 * tc[1] - a[2] / 6.0 * tc[2] - a[3] / 12.0 * tc[3] -
               a[4] / 20.0 * tc[4] + a[5] * invT - a[6];
    }
}

// sensible e/(R T), Source/CHEM/LiDryer.H:531-614
void species_internal_energy(double* u, const double* tc)
{
    double T = tc[1];
    double invT = 1.0 / T;
    for (int s = 0; s < NS; ++s) {
        if (T < 1000) {
            const double* a = nasa_lo[s];
            u[s] = a[0] - 1.0 + a[1] / 2.0 * tc[1] + a[2] / 3.0 * tc[2] + a[3] / 4.0 * tc[3] + a[4] / 5.0 * tc[4];
        } else {
            const double* a = nasa_hi[s];
            u[s] = a[0] - 1.0 + a[1] / 2.0 * tc[1] + a[2] / 3.0 * tc[2] + a[3] / 4.0 * tc[3] + a[4] / 5.0 * tc[4] +
                   (nasa_hi[s][5] - nasa_lo[s][5]) * invT;
        }
    }
}

// sensible h/R in K, Source/CHEM/LiDryer.H:616-695
void species_enthalpy(double T, double* h)
{
    double tc[] = {T, T * T, T * T * T, T * T * T * T, T * T * T * T * T};
    for (int s = 0; s < NS; ++s) {
        if (T < 1000) {
            const double* a = nasa_lo[s];
            h[s] = a[0] * tc[0] + a[1] / 2.0 * tc[1] + a[2] / 3.0 * tc[2] + a[3] / 4.0 * tc[3] + a[4] / 5.0 * tc[4];
        } else {
            const double* a = nasa_hi[s];
            h[s] = a[0] * tc[0] + a[1] / 2.0 * tc[1] + a[2] / 3.0 * tc[2] + a[3] / 4.0 * tc[3] + a[4] / 5.0 * tc[4] +
                   (nasa_hi[s][5] - nasa_lo[s][5]);
        }
    }
}

// cp/R, Source/CHEM/LiDryer.H:717-787
void cp_R(double* c, const double* tc)
{
    double T = tc[1];
    for (int s = 0; s < NS; ++s) {
        const double* a = T < 1000 ? nasa_lo[s] : nasa_hi[s];
        c[s] = a[0] + a[1] * tc[1] + a[2] * tc[2] + a[3] * tc[3] + a[4] * tc[4];
    }
}

// J/m^3, Source/CHEM/LiDryer.H:697-715
double CKUBMS(double T, const double* rhoi)
{
    double result = 0;
    double tc[] = {0, T, T * T, T * T * T, T * T * T * T};
    double ums[NS];
    double RT = Ru * T;
    species_internal_energy(ums, tc);
    for (int n = 0; n < NS; ++n) result += rhoi[n] * ums[n] / mw[n];
    return result * RT;
}

// J/(m^3 K), Source/CHEM/LiDryer.H:789-816
double CKCPBS(double T, const double* rhoi)
{
    double result = 0;
    double tc[] = {0, T, T * T, T * T * T, T * T * T * T};
    double cpor[NS];
    cp_R(cpor, tc);
    for (int n = 0; n < NS; ++n) result += cpor[n] * rhoi[n] / mw[n];
    return result * Ru;
}

double CKCVBS(double T, const double* rhoi)
{
    double result = 0;
    double tc[] = {0, T, T * T, T * T * T, T * T * T * T};
    double cpor[NS];
    cp_R(cpor, tc);
    for (int n = 0; n < NS; ++n) result += (cpor[n] - 1.0) * rhoi[n] / mw[n];
    return result * Ru;
}

// Source/CHEM/LiDryer.H:818-834
double CKGAMMA(double T, const double* rhoi)
{
    double cp = 0, cv = 0;
    double tc[] = {0, T, T * T, T * T * T, T * T * T * T};
    double cpor[NS];
    cp_R(cpor, tc);
    for (int n = 0; n < NS; ++n) {
        cp += cpor[n] * rhoi[n] / mw[n];
        cv += (cpor[n] - 1.0) * rhoi[n] / mw[n];
    }
    return cp / cv;
}

// Source/CHEM/LiDryer.H:10-66
double CKPY(const double* rhoi, double T)
{
    double YOW = 0;
    for (int i = 0; i < NS; ++i) YOW += rhoi[i] / mw[i];
    return Ru * T * YOW;
}

double CKRHOY(double P, double T, const double* y)
{
    double YOW = 0, tmp[NS];
    for (int i = 0; i < NS; ++i) tmp[i] = y[i] / mw[i];
    for (int i = 0; i < NS; ++i) YOW += tmp[i];
    return P / (Ru * T * YOW);
}

void CKYTX(const double* y, double* x)
{
    double YOW = 0, tmp[NS];
    for (int i = 0; i < NS; i++) tmp[i] = y[i] / mw[i];
    for (int i = 0; i < NS; ++i) YOW += tmp[i];
    double YOWINV = 1.0 / YOW;
    for (int i = 0; i < NS; i++) x[i] = y[i] / mw[i] * YOWINV;
}

void CKXTY(const double* x, double* y)
{
    double XW = 0.0;
    for (int i = 0; i < NS; ++i) XW += x[i] * mw[i];
    for (int i = 0; i < NS; ++i) y[i] = x[i] * mw[i] / XW;
}

// Newton inversion of the sensible energy, Source/CHEM/LiDryer.H:836-888
double T_given_ey(double e, const double* rhoi, int* iters)
{
    const int maxiter = 200;
    const double tol = 1e-7;
    double ein = e;
    double tmin = 250, tmax = 3500;
    double emin = CKUBMS(tmin, rhoi);
    double emax = CKUBMS(tmax, rhoi);
    if (iters) *iters = 0;
    if (ein < emin) return tmin - (emin - ein) / CKCVBS(tmin, rhoi);
    if (ein > emax) return tmax - (emax - ein) / CKCVBS(tmax, rhoi);
    double As = 0;
    for (int n = 0; n < NS; ++n) As += A0(n) * Ru / mw[n] * rhoi[n];
    double t1 = e / As;
    if (t1 < tmin || t1 > tmax) t1 = tmin + (tmax - tmin) / (emax - emin) * (ein - emin);
    int i;
    for (i = 0; i < maxiter; ++i) {
        double e1 = CKUBMS(t1, rhoi);
        double cv = CKCVBS(t1, rhoi);
        double dt = (ein - e1) / cv;
        if (dt > 100.0) dt = 100.0;
        else if (dt < -100.0) dt = -100.0;
        else if (std::fabs(dt) < tol) break;
        else if (t1 + dt == t1) break;
        t1 += dt;
    }
    if (iters) *iters = i;
    return t1;
}


// ---------------------------------------------------------------------------
// Transport: mixture-averaged viscosity, conductivity and diffusion
// coefficients, Source/CHEM/ChemViscous.H:8-106 with the fits of
// Source/Constants.cpp:84-125 (viscosity: sqrt(mu/sqrt(T)) as a quartic in
// ln T; conductivity: lambda/sqrt(T); binary diffusion: D_jk / T^1.5).
// ---------------------------------------------------------------------------
const double visc_fit[NS][5] = {
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
const double cond_fit[NS][5] = {
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
const double diff_fit[45][5] = {
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


inline double dot5(const double* x, const double* y)
{
    return x[0] * y[0] + x[1] * y[1] + x[2] * y[2] + x[3] * y[3] + x[4] * y[4];
}

// mu (Pa s) and lambda (W/m/K), Source/CHEM/ChemViscous.H:16-64
void mixture_properties(double T, const double* X, double& mu, double& lambda)
{
    double mui[NS], sqmui[NS], lambdai[NS];
    double phi[NS][NS] = {{0.0}};
    double sqT = std::sqrt(T), logT = std::log(T);
    double sqsqT = std::sqrt(sqT);
    double poly[5] = {1.0, logT, logT * logT, logT * logT * logT, logT * logT * logT * logT};
    for (int k = 0; k < NS; k++) {
        sqmui[k] = sqsqT * dot5(poly, visc_fit[k]);
        lambdai[k] = sqT * dot5(poly, cond_fit[k]);
        mui[k] = (sqmui[k] * sqmui[k]);
    }
    double sum1 = 0.0, sum2 = 0.0;
    for (int k = 0; k < NS; k++) {
        sum1 += X[k] * lambdai[k];
        sum2 += X[k] / lambdai[k];
    }
    lambda = 0.5 * (sum1 + 1.0 / sum2);
    // Wilke's mixing rule
    for (int k = 0; k < NS; k++) {
        for (int j = 0; j < k; j++) {
            double wratiojk = mw[j] / mw[k];
            double vratiokj = mui[k] / mui[j];
            double factor1 = 1.0 + (sqmui[k] / sqmui[j]) * std::sqrt(std::sqrt(wratiojk));
            double tmp = factor1 * factor1 / std::sqrt(8.0 + 8.0 * mw[k] / mw[j]);
            phi[k][j] = tmp;
            phi[j][k] = tmp / (vratiokj * wratiojk);
        }
        phi[k][k] = 1.0;
    }
    mu = 0;
    for (int k = 0; k < NS; k++) {
        double tmp = 0;
        for (int j = 0; j < NS; j++) tmp += X[j] * phi[k][j];
        mu += X[k] * mui[k] / tmp;
    }
}

// mixture-averaged diffusion coefficients (m^2/s), Source/CHEM/ChemViscous.H:66-106
void mix_diff_coeffs_mass(double T, double p, const double* X, double* D)
{
    double logT = std::log(T);
    double poly[5] = {1.0, logT, logT * logT, logT * logT * logT, logT * logT * logT * logT};
    double bdiff[NS][NS];
    int ic = 0;
    for (int n = 0; n < NS; n++)
        for (int nn = n; nn < NS; nn++) {
            bdiff[n][nn] = T * std::sqrt(T) * dot5(poly, diff_fit[ic]);
            bdiff[nn][n] = bdiff[n][nn];
            ic++;
        }
    double mmw = 0;
    for (int n = 0; n < NS; ++n) mmw += X[n] * mw[n];
    for (int n = 0; n < NS; n++) {
        double sum1 = 0.0, sum2 = 0.0;
        for (int nn = 0; nn < NS; nn++) {
            if (nn == n) continue;
            sum1 += X[nn] / bdiff[n][nn];
            sum2 += X[nn] * mw[nn] / bdiff[n][nn];
        }
        sum1 *= p;
        sum2 *= p * X[n] / (mmw - mw[n] * X[n]);
        D[n] = 1.0 / (sum1 + sum2);
    }
}

// Arrhenius forward rate constants (CGS), Source/CHEM/LiDryer.H:171-191:
// form 0: A exp(beta lnT - TA/T); 1: A exp(beta lnT); 2: A exp(c/T); 3: A
struct Arr { double A, beta, c; int form; };
const Arr kf_tab[21] = {
    {3.55e+15, -0.41, 8353.444256079512, 0}, {50800.0, 2.67, 3165.2508657072367, 0},
    {2.16e+08, 1.51, 1726.0429999007667, 0}, {2.97e+06, 2.02, 6743.141748883461, 0},
    {4.58e+19, -1.4, 52526.05490660117, 0}, {6.16e+15, -0.5, 0, 1}, {4.71e+18, -1, 0, 1}, {3.8e+22, -2, 0, 1},
    {1.48e+12, 0.6, 0, 1}, {1.66e+13, 0, -412.64001746898793, 2}, {7.08e+13, 0, -150.96586004962973, 2},
    {3.25e+13, 0, 0, 3}, {2.89e+13, 0, +251.60976674938289, 2}, {4.2e+14, 0, -6028.570011315214, 2},
    {1.3e+11, 0, +820.2478396029883, 2}, {2.95e+14, 0, -24355.825421340265, 2}, {2.41e+13, 0, -1997.7815479901, 2},
    {4.82e+13, 0, -4000.595291315188, 2}, {9.55e+06, 2, 1997.7815479901, 0}, {1.0e+12, 0, 0, 3},
    {5.8e+14, 0, -4810.778740248201, 2}};

// Equilibrium constants, Source/CHEM/LiDryer.H:197-219: exp((g[a0]+g[a1]) - (g[b0]+g[b1]))
// times 1 (fac 0), 1/RsT (fac 1) or RsT (fac 2); -1 = absent term
struct KcDef { int a0, a1, b0, b1, fac; };
const KcDef kc_tab[21] = {
    {3, 1, 4, 5, 0}, {4, 0, 3, 5, 0}, {0, 5, 2, 3, 0}, {4, 2, 5, 5, 0}, {0, -1, 3, 3, 1}, {4, 4, 1, -1, 2},
    {4, 3, 5, -1, 2}, {3, 5, 2, -1, 2}, {3, 1, 6, -1, 2}, {6, 3, 0, 1, 0}, {6, 3, 5, 5, 0}, {6, 4, 1, 5, 0},
    {6, 5, 2, 1, 0}, {6, 6, 7, 1, 0}, {6, 6, 7, 1, 0}, {7, -1, 5, 5, 1}, {7, 3, 2, 5, 0}, {7, 3, 6, 0, 0},
    {7, 4, 5, 6, 0}, {7, 5, 6, 2, 0}, {7, 5, 6, 2, 0}};

// Reactions, Source/CHEM/LiDryer.H:232-510.  f0,f1 / r0,r1: species of the
// forward / reverse concentration products (-1 = absent); kf: rate index;
// dup: second rate index of a duplicate (-1 none); tb: third body
// (0 none, 1 M with H2 2.5 H2O 12, 2 falloff HO2 (Fcent 0.8), 3 falloff H2O2 (Fcent 0.5)).
struct Reac { int f0, f1, r0, r1, kf, dup, tb; int nuf[NS]; int nur[NS]; };
const Reac rx[NR] = {
    {3, 1, 4, 5, 0, -1, 0, {0, 1, 0, 1, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 1, 1, 0, 0, 0}},
    {4, 0, 3, 5, 1, -1, 0, {1, 0, 0, 0, 1, 0, 0, 0, 0}, {0, 0, 0, 1, 0, 1, 0, 0, 0}},
    {0, 5, 2, 3, 2, -1, 0, {1, 0, 0, 0, 0, 1, 0, 0, 0}, {0, 0, 1, 1, 0, 0, 0, 0, 0}},
    {4, 2, 5, 5, 3, -1, 0, {0, 0, 1, 0, 1, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 2, 0, 0, 0}},
    {0, -1, 3, 3, 4, -1, 1, {1, 0, 0, 0, 0, 0, 0, 0, 0}, {0, 0, 0, 2, 0, 0, 0, 0, 0}},
    {4, 4, 1, -1, 5, -1, 1, {0, 0, 0, 0, 2, 0, 0, 0, 0}, {0, 1, 0, 0, 0, 0, 0, 0, 0}},
    {4, 3, 5, -1, 6, -1, 1, {0, 0, 0, 1, 1, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 1, 0, 0, 0}},
    {3, 5, 2, -1, 7, -1, 1, {0, 0, 0, 1, 0, 1, 0, 0, 0}, {0, 0, 1, 0, 0, 0, 0, 0, 0}},
    {3, 1, 6, -1, 8, -1, 2, {0, 1, 0, 1, 0, 0, 0, 0, 0}, {0, 0, 0, 0, 0, 0, 1, 0, 0}},
    {6, 3, 0, 1, 9, -1, 0, {0, 0, 0, 1, 0, 0, 1, 0, 0}, {1, 1, 0, 0, 0, 0, 0, 0, 0}},
    {6, 3, 5, 5, 10, -1, 0, {0, 0, 0, 1, 0, 0, 1, 0, 0}, {0, 0, 0, 0, 0, 2, 0, 0, 0}},
    {6, 4, 1, 5, 11, -1, 0, {0, 0, 0, 0, 1, 0, 1, 0, 0}, {0, 1, 0, 0, 0, 1, 0, 0, 0}},
    {6, 5, 2, 1, 12, -1, 0, {0, 0, 0, 0, 0, 1, 1, 0, 0}, {0, 1, 1, 0, 0, 0, 0, 0, 0}},
    {6, 6, 7, 1, 13, 14, 0, {0, 0, 0, 0, 0, 0, 2, 0, 0}, {0, 1, 0, 0, 0, 0, 0, 1, 0}},
    {7, -1, 5, 5, 15, -1, 3, {0, 0, 0, 0, 0, 0, 0, 1, 0}, {0, 0, 0, 0, 0, 2, 0, 0, 0}},
    {7, 3, 2, 5, 16, -1, 0, {0, 0, 0, 1, 0, 0, 0, 1, 0}, {0, 0, 1, 0, 0, 1, 0, 0, 0}},
    {7, 3, 6, 0, 17, -1, 0, {0, 0, 0, 1, 0, 0, 0, 1, 0}, {1, 0, 0, 0, 0, 0, 1, 0, 0}},
    {7, 4, 5, 6, 18, -1, 0, {0, 0, 0, 0, 1, 0, 0, 1, 0}, {0, 0, 0, 0, 0, 1, 1, 0, 0}},
    {7, 5, 6, 2, 19, 20, 0, {0, 0, 0, 0, 0, 1, 0, 1, 0}, {0, 0, 1, 0, 0, 0, 1, 0, 0}}};

inline double prod(const double* sc, int a, int b) { return b < 0 ? sc[a] : sc[a] * sc[b]; }

// Production rates (mol/cm^3/s) and their Jacobian d wdot_l / d c_n at fixed T,
// Source/CHEM/LiDryer.H:156-529.
void production_rate(double* wdot, double (*Arate)[NS], const double* sc, double T)
{
    const double eps = 1e-20;
    double tc[5];
    tc[0] = std::log(T);
    tc[1] = T;
    tc[2] = T * (T);
    tc[3] = tc[2] * tc[1];
    tc[4] = tc[2] * tc[2];
    double invT = 1.0 / T;
    double kfs[21];
    for (int r = 0; r < 21; ++r) {
        const Arr& a = kf_tab[r];
        if (a.form == 0) kfs[r] = a.A * std::exp(a.beta * tc[0] - a.c * invT);
        else if (a.form == 1) kfs[r] = a.A * std::exp(a.beta * tc[0]);
        else if (a.form == 2) kfs[r] = a.A * std::exp(a.c * invT);
        else kfs[r] = a.A;
    }
    double g[NS];
    gibbs(g, tc);
    double RsT = Ru / Patm * 1e6 * T;
    double Kc[21];
    for (int r = 0; r < 21; ++r) {
        const KcDef& d = kc_tab[r];
        double sa = d.a1 < 0 ? g[d.a0] : g[d.a0] + g[d.a1];
        double sb = d.b1 < 0 ? g[d.b0] : g[d.b0] + g[d.b1];
        double e = std::exp((sa) - (sb));
        Kc[r] = d.fac == 0 ? e : (d.fac == 1 ? 1.0 / RsT * e : RsT * e);
    }
    double mixture = 0.0;
    for (int n = 0; n < NS; n++) {
        mixture += sc[n];
        wdot[n] = 0.0;
        for (int l = 0; l < NS; l++) Arate[n][l] = 0.0;
    }
    double q_f[NR], q_r[NR];
    for (int m = 0; m < NR; ++m) {
        const Reac& R = rx[m];
        double phi_f = prod(sc, R.f0, R.f1);
        double phi_r = prod(sc, R.r0, R.r1);
        double k_f = kfs[R.kf];
        if (R.tb == 1) {
            double alpha = mixture + 1.5 * sc[0] + 11 * sc[2];
            k_f = alpha * k_f;
        } else if (R.tb >= 2) {
            double alpha, redP, Fc;
            if (R.tb == 2) {
                alpha = mixture + sc[0] + 10 * sc[2] + -0.22 * sc[1];
                redP = alpha / k_f * 6.37e+20 * std::exp(-1.72 * tc[0] - 261.6741574193582 * invT);
                Fc = 0.8;
            } else {
                alpha = mixture + 1.5 * sc[0] + 11 * sc[2];
                redP = alpha / k_f * 1.20e+17 * std::exp(-22846.166820843966 * invT);
                Fc = 0.5;
            }
            double logPred = std::log10(redP);
            double logFcent = std::log10(Fc);
            double troe_c = -0.4 - 0.67 * logFcent;
            double troe_n = 0.75 - 1.27 * logFcent;
            double troe = (troe_c + logPred) / (troe_n - 0.14 * (troe_c + logPred));
            double F_troe = std::pow(10.0, logFcent / (1.0 + troe * troe));
            k_f *= (redP / (1.0 + redP)) * F_troe;
        }
        q_f[m] = phi_f * k_f;
        q_r[m] = phi_r * (k_f / Kc[R.kf]);
        if (R.dup >= 0) {
            double k2 = kfs[R.dup];
            q_f[m] += phi_f * k2;
            q_r[m] += phi_r * (k2 / Kc[R.dup]);
        }
    }
    for (int m = 0; m < NR; ++m) {
        const Reac& R = rx[m];
        double wf1 = q_f[m], wr1 = q_r[m];
        for (int n = 0; n < NS; ++n) wdot[n] += (wf1 - wr1) * (R.nur[n] - R.nuf[n]);
        for (int n = 0; n < NS; ++n) {
            double Awf = R.nuf[n] * wf1 / (sc[n] + eps);
            double Awr = R.nur[n] * wr1 / (sc[n] + eps);
            for (int l = 0; l < NS; ++l) Arate[l][n] += (Awf - Awr) * (R.nur[l] - R.nuf[l]);
        }
    }
}

// 9x9 Gaussian elimination with partial pivoting, Source/CHEM/LiDryer.H:890-942
void gauss(double* x, const double (*A)[NS], const double* b)
{
    double U[NS][NS + 1];
    for (int i = 0; i < NS; i++) {
        U[i][NS] = b[i];
        for (int j = 0; j < NS; j++) U[i][j] = A[i][j];
    }
    for (int k = 0; k < NS; k++) {
        int p = k;
        double maxPivot = 0.0;
        for (int i = k; i < NS; i++)
            if (std::fabs(U[i][k]) > maxPivot) { maxPivot = std::fabs(U[i][k]); p = i; }
        if (p != k)
            for (int i = k; i < NS + 1; i++) std::swap(U[p][i], U[k][i]);
        for (int i = k + 1; i < NS; i++) {
            double m = U[i][k] / U[k][k];
            for (int j = k; j < NS + 1; j++) U[i][j] -= m * U[k][j];
        }
    }
    for (int k = NS - 1; k >= 0; k--) {
        double sum = U[k][NS];
        for (int j = k + 1; j < NS; j++) sum -= U[k][j] * x[j];
        x[k] = sum / U[k][k];
    }
}

// One linearised-implicit chemistry step of one cell (Source/CHEM/ChemAdvance.cpp:36-81):
// u[5] conserved state (E sensible + kinetic), rhoi[9] partial densities, in place.
void chem_step(double* u, double* rhoi, double dt, double smallr)
{
    double rho = std::max(u[0], smallr);
    double rhoinv = 1.0 / rho;
    double ux = u[1] * rhoinv, uy = u[2] * rhoinv, uz = u[3] * rhoinv;
    double kineng = 0.5 * (ux * ux + uy * uy + uz * uz);
    double ei = u[4] - rho * kineng;
    double rhoi_0[NS], c[NS], wdot[NS], Arate[NS][NS], A1[NS][NS], rhs[NS], drho[NS];
    for (int n = 0; n < NS; ++n) rhoi_0[n] = rhoi[n];
    double T = T_given_ey(ei, rhoi_0, nullptr);
    for (int n = 0; n < NS; n++) c[n] = rhoi_0[n] / mw[n] * 1e-6;
    production_rate(wdot, Arate, c, T);
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) {
            A1[i][j] = (i == j ? 1.0 : 0.0);
            A1[i][j] -= Arate[i][j] * mw[i] / mw[j] * dt;
        }
    for (int n = 0; n < NS; ++n) rhs[n] = wdot[n] * mw[n] * 1e6 * dt;
    gauss(drho, A1, rhs);
    for (int n = 0; n < NS; ++n) {
        rhoi[n] += drho[n];
        if (rhoi[n] < 0) rhoi[n] = 0.0;
    }
    double dei = 0;
    for (int n = 0; n < NS; ++n) dei += -HP(n) * Ru / mw[n] * drho[n];
    u[4] += dei;
}

// Species clip + rescale to rho, Source/EBR.cpp:341-355
void renorm(const double* u, double* rhoi)
{
    double rho0 = 0;
    for (int n = 0; n < NS; ++n) {
        if (rhoi[n] < 0) rhoi[n] = 0.0;
        rho0 += rhoi[n];
    }
    double tmp = u[0] / rho0;
    for (int n = 0; n < NS; ++n) rhoi[n] *= tmp;
}

// Adaptive integration of one constant-volume reactor with step doubling on
// the linearised-implicit step: one step of h (y1) against two of h/2 (y2);
// error max_k |y2-y1| / (atol*rho + rtol*|y2|) over the partial densities; on
// acceptance the Richardson value 2*y2 - y1 (second order, densities clipped at
// 0) is kept; h <- h * clamp(0.9/sqrt(err), 0.2, 4).  Returns accepted steps,
// or -1 when max_steps attempts did not reach t_end.
int reactor_adaptive(double* u, double* rhoi, double t_end, double h0, double rtol, double atol, int max_steps,
                     int* rejects)
{
    double t = 0.0, h = h0;
    int acc = 0, rej = 0;
    const double rho = u[0];
    while (t < t_end) {
        if (acc + rej >= max_steps) { if (rejects) *rejects = rej; return -1; }
        if (t + h > t_end) h = t_end - t;
        double u1[5], r1[NS], u2[5], r2[NS];
        std::memcpy(u1, u, sizeof(u1)); std::memcpy(r1, rhoi, sizeof(r1));
        std::memcpy(u2, u, sizeof(u2)); std::memcpy(r2, rhoi, sizeof(r2));
        chem_step(u1, r1, h, 1e-19);
        chem_step(u2, r2, 0.5 * h, 1e-19);
        chem_step(u2, r2, 0.5 * h, 1e-19);
        double err = 0.0;
        for (int n = 0; n < NS; ++n) {
            double e = std::fabs(r2[n] - r1[n]) / (atol * rho + rtol * std::fabs(r2[n]));
            err = std::max(err, e);
        }
        if (err <= 1.0) {
            for (int n = 0; n < NS; ++n) rhoi[n] = std::max(2.0 * r2[n] - r1[n], 0.0);
            u[4] = 2.0 * u2[4] - u1[4];
            t += h;
            ++acc;
            double f = err > 0.0 ? 0.9 / std::sqrt(err) : 4.0;
            h *= std::min(4.0, std::max(0.2, f));
        } else {
            ++rej;
            h *= std::max(0.2, 0.9 / std::sqrt(err));
        }
    }
    if (rejects) *rejects = rej;
    return acc;
}

// ---------------------------------------------------------------------------
// Rosenbrock integration of one constant-volume reactor (no counterpart in the
// reference, whose chemistry takes fixed linearised-implicit steps): the ODE
// d rho_k/dt = W_k wdot_k(c, T) with T from the sensible energy E(rho) =
// E0 - sum_k hf_k (rho_k - rho_k0) (the energy bookkeeping of
// Source/CHEM/ChemAdvance.cpp:48-81), by RODAS3 (Sandu, Verwer, Blom,
// Spee, Carmichael, Potra, Atmos. Environ. 31 (1997) 3459-3472): 4 stages,
// order 3 with an embedded order-2 solution, L-stable and stiffly accurate,
// in the transformed form
//   (I/(h gamma) - J) K_i = f(y + sum_j a_ij K_j) + sum_j (c_ij/h) K_j,
//   y1 = y + sum_i m_i K_i,  error estimate K_4.
// J is the full Jacobian: the mass-action part at fixed T plus
// (d f/d T)(d T/d rho), d f/d T by a forward difference of 1e-7 T and
// d T/d rho_j = (-hf_j - e_j(T)) / cv from the energy relation.
// ---------------------------------------------------------------------------
const double kRosGamma = 0.5;

// Troe fall-off factor P(Pr) = Pr/(1+Pr) F(Pr) of production_rate and dP/dPr
void troe_and_slope(double redP, double Fc, double& P, double& dP)
{
    const double logPred = std::log10(redP);
    const double logFcent = std::log10(Fc);
    const double troe_c = -0.4 - 0.67 * logFcent;
    const double troe_n = 0.75 - 1.27 * logFcent;
    const double uu = troe_c + logPred;
    const double den = troe_n - 0.14 * uu;
    const double t = uu / den;
    const double F = std::pow(10.0, logFcent / (1.0 + t * t));
    P = (redP / (1.0 + redP)) * F;
    const double q = 1.0 + t * t;
    dP = F / ((1.0 + redP) * (1.0 + redP)) +
         F / (1.0 + redP) * (-2.0 * t * logFcent * troe_n) / (q * q * den * den);
}

// The complete d wdot / d c (CGS) at fixed T: mass action by the product
// rule (exact where a species is absent), the third-body efficiencies and
// the Troe fall-off, with the rates of production_rate
void jacobian_fixed_T(const double* sc, double T, double (*J)[NS])
{
    double tc[5];
    tc[0] = std::log(T);
    tc[1] = T;
    tc[2] = T * (T);
    tc[3] = tc[2] * tc[1];
    tc[4] = tc[2] * tc[2];
    const double invT = 1.0 / T;
    double kfs[21];
    for (int r = 0; r < 21; ++r) {
        const Arr& a = kf_tab[r];
        if (a.form == 0) kfs[r] = a.A * std::exp(a.beta * tc[0] - a.c * invT);
        else if (a.form == 1) kfs[r] = a.A * std::exp(a.beta * tc[0]);
        else if (a.form == 2) kfs[r] = a.A * std::exp(a.c * invT);
        else kfs[r] = a.A;
    }
    double g[NS];
    gibbs(g, tc);
    const double RsT = Ru / Patm * 1e6 * T;
    double Kc[21];
    for (int r = 0; r < 21; ++r) {
        const KcDef& d = kc_tab[r];
        double sa = d.a1 < 0 ? g[d.a0] : g[d.a0] + g[d.a1];
        double sb = d.b1 < 0 ? g[d.b0] : g[d.b0] + g[d.b1];
        double e = std::exp((sa) - (sb));
        Kc[r] = d.fac == 0 ? e : (d.fac == 1 ? 1.0 / RsT * e : RsT * e);
    }
    double mixture = 0.0;
    for (int n = 0; n < NS; ++n) {
        mixture += sc[n];
        for (int l = 0; l < NS; ++l) J[n][l] = 0.0;
    }
    for (int m = 0; m < NR; ++m) {
        const Reac& R = rx[m];
        const double phi_f = prod(sc, R.f0, R.f1), phi_r = prod(sc, R.r0, R.r1);
        double dpf[NS] = {0}, dpr[NS] = {0};
        if (R.f1 < 0) dpf[R.f0] += 1.0;
        else { dpf[R.f0] += sc[R.f1]; dpf[R.f1] += sc[R.f0]; }
        if (R.r1 < 0) dpr[R.r0] += 1.0;
        else { dpr[R.r0] += sc[R.r1]; dpr[R.r1] += sc[R.r0]; }
        const double k0 = kfs[R.kf];
        double keff = k0, dk[NS] = {0};
        if (R.tb != 0) {
            double eff[NS];
            for (int n = 0; n < NS; ++n) eff[n] = 1.0;
            double alpha;
            if (R.tb == 2) {
                eff[0] += 1.0; eff[2] += 10.0; eff[1] += -0.22;
                alpha = mixture + sc[0] + 10 * sc[2] + -0.22 * sc[1];
            } else {
                eff[0] += 1.5; eff[2] += 11.0;
                alpha = mixture + 1.5 * sc[0] + 11 * sc[2];
            }
            if (R.tb == 1) {
                keff = alpha * k0;
                for (int n = 0; n < NS; ++n) dk[n] = eff[n] * k0;
            } else {
                const double kl = R.tb == 2 ? 6.37e+20 * std::exp(-1.72 * tc[0] - 261.6741574193582 * invT)
                                            : 1.20e+17 * std::exp(-22846.166820843966 * invT);
                const double redP = alpha / k0 * kl;
                double P, dP;
                troe_and_slope(redP, R.tb == 2 ? 0.8 : 0.5, P, dP);
                keff = k0 * P;
                for (int n = 0; n < NS; ++n) dk[n] = dP * kl * eff[n];
            }
        }
        const double diff = phi_f - phi_r / Kc[R.kf];
        double dq[NS];
        for (int n = 0; n < NS; ++n) dq[n] = keff * (dpf[n] - dpr[n] / Kc[R.kf]) + dk[n] * diff;
        if (R.dup >= 0) {
            const double k2 = kfs[R.dup];
            for (int n = 0; n < NS; ++n) dq[n] += k2 * (dpf[n] - dpr[n] / Kc[R.dup]);
        }
        for (int l = 0; l < NS; ++l) {
            const int nu = R.nur[l] - R.nuf[l];
            if (nu == 0) continue;
            for (int n = 0; n < NS; ++n) J[l][n] += nu * dq[n];
        }
    }
}

// Newton inversion of the sensible energy from a temperature guess (the
// iteration of T_given_ey), T_given_ey itself when the guess or an iterate
// leaves its table range
double T_newton(double e, const double* rhoi, double Tg)
{
    const double tol = 1e-7;
    if (!(Tg >= 250.0 && Tg <= 3500.0)) return T_given_ey(e, rhoi, nullptr);
    double t1 = Tg;
    for (int i = 0; i < 50; ++i) {
        double e1 = CKUBMS(t1, rhoi);
        double cv = CKCVBS(t1, rhoi);
        double dt = (e - e1) / cv;
        if (dt > 100.0) dt = 100.0;
        else if (dt < -100.0) dt = -100.0;
        else if (std::fabs(dt) < tol) return t1;
        else if (t1 + dt == t1) return t1;
        t1 += dt;
        if (t1 < 250.0 || t1 > 3500.0) return T_given_ey(e, rhoi, nullptr);
    }
    return t1;
}

// f = d rho/dt (kg/m^3/s) at T (production_rate); J = d f/d rho at fixed T
// (jacobian_fixed_T) when not null
void reactor_f(const double* y, double T, double* f, double (*J)[NS])
{
    double c[NS], wdot[NS], Ar[NS][NS];
    for (int n = 0; n < NS; ++n) c[n] = y[n] / mw[n] * 1e-6;
    production_rate(wdot, Ar, c, T);
    for (int i = 0; i < NS; ++i) f[i] = wdot[i] * mw[i] * 1e6;
    if (J) {
        double Jc[NS][NS];
        jacobian_fixed_T(c, T, Jc);
        for (int i = 0; i < NS; ++i)
            for (int j = 0; j < NS; ++j) J[i][j] = Jc[i][j] * mw[i] / mw[j];
    }
}

// LU with partial pivoting (first maximum), row exchanges in piv.  A row
// exchange at step k moves columns k.. only: the multipliers of the earlier
// columns stay with the rows they were applied to, which is the order in
// which lu_solve replays exchanges and eliminations.
void lu_factor(double (*M)[NS], int* piv)
{
    for (int k = 0; k < NS; ++k) {
        int p = k;
        double mx = 0.0;
        for (int i = k; i < NS; ++i)
            if (std::fabs(M[i][k]) > mx) { mx = std::fabs(M[i][k]); p = i; }
        piv[k] = p;
        if (p != k)
            for (int j = k; j < NS; ++j) std::swap(M[p][j], M[k][j]);
        for (int i = k + 1; i < NS; ++i) {
            const double m = M[i][k] / M[k][k];
            M[i][k] = m;
            for (int j = k + 1; j < NS; ++j) M[i][j] -= m * M[k][j];
        }
    }
}

void lu_solve(const double (*M)[NS], const int* piv, double* b)
{
    for (int k = 0; k < NS; ++k) {
        if (piv[k] != k) std::swap(b[k], b[piv[k]]);
        for (int i = k + 1; i < NS; ++i) b[i] -= M[i][k] * b[k];
    }
    for (int k = NS - 1; k >= 0; --k) {
        double sum = b[k];
        for (int j = k + 1; j < NS; ++j) sum -= M[k][j] * b[j];
        b[k] = sum / M[k][k];
    }
}

// energy change of a change dy of the partial densities (formation enthalpies)
double energy_of(const double* dy)
{
    double de = 0.0;
    for (int n = 0; n < NS; ++n) de += -HP(n) * Ru / mw[n] * dy[n];
    return de;
}

// One RODAS3 step of h from (y, E) with temperature guess Tg: the new state
// in ynew/Enew, the temperature of the last stage in Tlast; returns the
// error max_k |K4_k| / (atol rho + rtol max(|y_k|, |ynew_k|)).
double rodas3_step(const double* y, double E, double Tg, double h, double rho, double rtol, double atol, double* ynew,
                   double& Enew, double& Tlast)
{
    const double T0 = T_newton(E, y, Tg);
    double f0[NS], J[NS][NS], f1[NS];
    reactor_f(y, T0, f0, J);
    const double dT = 1e-7 * T0;
    reactor_f(y, T0 + dT, f1, nullptr);
    const double tc[5] = {0.0, T0, T0 * T0, T0 * T0 * T0, T0 * T0 * T0 * T0};
    double ue[NS];
    species_internal_energy(ue, tc);
    const double cv = CKCVBS(T0, y);
    double dTdy[NS], dfdT[NS];
    for (int j = 0; j < NS; ++j) dTdy[j] = (-HP(j) * Ru / mw[j] - ue[j] * Ru * T0 / mw[j]) / cv;
    for (int i = 0; i < NS; ++i) dfdT[i] = (f1[i] - f0[i]) / dT;
    double M[NS][NS];
    const double gh = 1.0 / (h * kRosGamma);
    for (int i = 0; i < NS; ++i)
        for (int j = 0; j < NS; ++j) M[i][j] = -(J[i][j] + dfdT[i] * dTdy[j]) + (i == j ? gh : 0.0);
    int piv[NS];
    lu_factor(M, piv);
    const double hinv = 1.0 / h;
    double K1[NS], K2[NS], K3[NS], K4[NS], ys[NS], fs[NS], dy[NS];
    for (int n = 0; n < NS; ++n) K1[n] = f0[n];
    lu_solve(M, piv, K1);
    for (int n = 0; n < NS; ++n) K2[n] = f0[n] + (4.0 * hinv) * K1[n];
    lu_solve(M, piv, K2);
    // stage 3 at y + 2 K1
    for (int n = 0; n < NS; ++n) { dy[n] = 2.0 * K1[n]; ys[n] = y[n] + dy[n]; }
    const double T3 = T_newton(E + energy_of(dy), ys, T0);
    reactor_f(ys, T3, fs, nullptr);
    for (int n = 0; n < NS; ++n) K3[n] = fs[n] + hinv * K1[n] - hinv * K2[n];
    lu_solve(M, piv, K3);
    // stage 4 at y + 2 K1 + K3
    for (int n = 0; n < NS; ++n) { dy[n] = 2.0 * K1[n] + K3[n]; ys[n] = y[n] + dy[n]; }
    const double T4 = T_newton(E + energy_of(dy), ys, T3);
    reactor_f(ys, T4, fs, nullptr);
    for (int n = 0; n < NS; ++n) K4[n] = fs[n] + hinv * K1[n] - hinv * K2[n] - (8.0 / 3.0 * hinv) * K3[n];
    lu_solve(M, piv, K4);
    double err = 0.0;
    bool finite = true;
    for (int n = 0; n < NS; ++n) {
        dy[n] = 2.0 * K1[n] + K3[n] + K4[n];
        ynew[n] = y[n] + dy[n];
        const double e = std::fabs(K4[n]) / (atol * rho + rtol * std::max(std::fabs(y[n]), std::fabs(ynew[n])));
        finite = finite && std::isfinite(e) && std::isfinite(ynew[n]);
        err = std::max(err, e);
    }
    Enew = E + energy_of(dy);
    Tlast = T4;
    // a stage that left the thermodynamic tables (a non-physical temperature,
    // non-finite rates) fails the step: it is retried with a smaller h
    return finite ? err : HUGE_VAL;
}

// Adaptive RODAS3 to t_end: on acceptance the densities are clipped at 0 (the
// energy follows the clipped change), h <- h clamp(0.9 err^(-1/3), 0.2, 6), not
// above h after a rejection; on rejection h <- h max(0.2, 0.9 err^(-1/3)).
// Returns accepted steps, or -1 when max_steps attempts did not reach t_end.
int reactor_rosenbrock(double* u, double* rhoi, double t_end, double h0, double rtol, double atol, int max_steps,
                       int* rejects)
{
    double t = 0.0, h = h0, E = u[4], Tg = 0.0;
    int acc = 0, rej = 0;
    bool last_rej = false;
    const double rho = u[0];
    while (t < t_end) {
        if (acc + rej >= max_steps) { if (rejects) *rejects = rej; return -1; }
        if (t + h > t_end) h = t_end - t;
        double yn[NS], En, Tl;
        const double err = rodas3_step(rhoi, E, Tg, h, rho, rtol, atol, yn, En, Tl);
        const double fac = 0.9 / std::cbrt(err);
        if (err <= 1.0) {
            double dy[NS];
            for (int n = 0; n < NS; ++n) {
                const double v = std::max(yn[n], 0.0);
                dy[n] = v - rhoi[n];
                rhoi[n] = v;
            }
            E += energy_of(dy);
            Tg = Tl;
            t += h;
            ++acc;
            double f = std::min(6.0, std::max(0.2, fac));
            if (last_rej) f = std::min(f, 1.0);
            h *= f;
            last_rej = false;
        } else {
            ++rej;
            h *= std::max(0.2, fac);
            last_rej = true;
        }
    }
    u[4] = E;
    if (rejects) *rejects = rej;
    return acc;
}

} // namespace

extern "C" {

// Rosenbrock (RODAS3) integration to t_end; steps[m] = accepted steps (-1 on failure)
void orc_reactor_rosenbrock(long n, double* u, double* rhoi, double t_end, double h0, double rtol, double atol,
                            int max_steps, int* steps, int* rejects)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (long m = 0; m < n; ++m)
        steps[m] = reactor_rosenbrock(u + 5 * m, rhoi + NS * m, t_end, h0, rtol, atol, max_steps,
                                      rejects ? rejects + m : nullptr);
}

// nsteps RODAS3 steps of h without error control (order checks); u, rhoi in place
void orc_rodas3_fixed(double* u, double* rhoi, double h, int nsteps)
{
    double E = u[4], Tg = 0.0;
    for (int s = 0; s < nsteps; ++s) {
        double yn[NS], En, Tl;
        rodas3_step(rhoi, E, Tg, h, u[0], 1e-4, 1e-9, yn, En, Tl);
        std::memcpy(rhoi, yn, sizeof(yn));
        E = En;
        Tg = Tl;
    }
    u[4] = E;
}

// the complete d wdot/d c at fixed T (CGS), row-major 9x9
void orc_jacobian_fixed_T(const double* c, double T, double* J)
{
    double Jc[NS][NS];
    jacobian_fixed_T(c, T, Jc);
    std::memcpy(J, Jc, sizeof(Jc));
}

// the reactor ODE's right-hand side at densities y, the energy following from
// (E0, y0): f = d rho/dt; *T = the temperature at y
void orc_reactor_rhs(double E0, const double* y0, const double* y, double* f, double* T)
{
    double dy[NS];
    for (int n = 0; n < NS; ++n) dy[n] = y[n] - y0[n];
    const double Ty = T_given_ey(E0 + energy_of(dy), y, nullptr);
    reactor_f(y, Ty, f, nullptr);
    if (T) *T = Ty;
}

void orc_chem_tables(double* mw_out, double* lo_out, double* hi_out)
{
    for (int n = 0; n < NS; ++n) {
        mw_out[n] = mw[n];
        for (int c = 0; c < 7; ++c) { lo_out[7 * n + c] = nasa_lo[n][c]; hi_out[7 * n + c] = nasa_hi[n][c]; }
    }
}

void orc_production_rate(const double* c, double T, double* wdot, double* jac)
{
    double A[NS][NS];
    production_rate(wdot, A, c, T);
    std::memcpy(jac, A, sizeof(A));
}

void orc_gauss(const double* A, const double* b, double* x)
{
    gauss(x, (const double(*)[NS])A, b);
}

double orc_T_given_ey(double e, const double* rhoi, int* iters) { return T_given_ey(e, rhoi, iters); }
double orc_CKUBMS(double T, const double* rhoi) { return CKUBMS(T, rhoi); }
double orc_CKCVBS(double T, const double* rhoi) { return CKCVBS(T, rhoi); }
double orc_CKCPBS(double T, const double* rhoi) { return CKCPBS(T, rhoi); }
double orc_CKGAMMA(double T, const double* rhoi) { return CKGAMMA(T, rhoi); }
double orc_CKPY(const double* rhoi, double T) { return CKPY(rhoi, T); }
double orc_CKRHOY(double P, double T, const double* y) { return CKRHOY(P, T, y); }
void orc_CKXTY(const double* x, double* y) { CKXTY(x, y); }
void orc_CKYTX(const double* y, double* x) { CKYTX(y, x); }
void orc_species_enthalpy(double T, double* h) { species_enthalpy(T, h); }
void orc_mixture_properties(double T, const double* X, double* mu, double* lambda) { mixture_properties(T, X, *mu, *lambda); }
void orc_mix_diff_coeffs_mass(double T, double p, const double* X, double* D) { mix_diff_coeffs_mass(T, p, X, D); }

// n cells: u (n x 5), rhoi (n x 9), in place, one step of dt each
void orc_chem_step_many(long n, double* u, double* rhoi, double dt, double smallr)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (long m = 0; m < n; ++m) chem_step(u + 5 * m, rhoi + NS * m, dt, smallr);
}

void orc_renorm_many(long n, const double* u, double* rhoi)
{
    for (long m = 0; m < n; ++m) renorm(u + 5 * m, rhoi + NS * m);
}

// Constant-volume H2/air reactor at (T0, p0, phi): X(H2):X(O2):X(N2) = 2phi : 1 : 3.76
void orc_reactor_init(long n, const double* T0, const double* p0, const double* phi, double* u, double* rhoi)
{
    for (long m = 0; m < n; ++m) {
        double X[NS] = {0}, Y[NS];
        double tot = 2.0 * phi[m] + 1.0 + 3.76;
        X[0] = 2.0 * phi[m] / tot;
        X[1] = 1.0 / tot;
        X[8] = 3.76 / tot;
        CKXTY(X, Y);
        double rho = CKRHOY(p0[m], T0[m], Y);
        double* r = rhoi + NS * m;
        for (int k = 0; k < NS; ++k) r[k] = Y[k] * rho;
        double* uu = u + 5 * m;
        uu[0] = rho; uu[1] = uu[2] = uu[3] = 0.0;
        uu[4] = CKUBMS(T0[m], r);
    }
}

// Fixed-step integration: nsteps linearised-implicit steps of dt
void orc_reactor_fixed(long n, double* u, double* rhoi, double dt, int nsteps)
{
#pragma omp parallel for schedule(dynamic, 16)
    for (long m = 0; m < n; ++m)
        for (int s = 0; s < nsteps; ++s) chem_step(u + 5 * m, rhoi + NS * m, dt, 1e-19);
}

// Adaptive integration to t_end; steps[m] = accepted steps (-1 on failure)
void orc_reactor_adaptive(long n, double* u, double* rhoi, double t_end, double h0, double rtol, double atol,
                          int max_steps, int* steps, int* rejects)
{
#pragma omp parallel for schedule(dynamic, 4)
    for (long m = 0; m < n; ++m)
        steps[m] = reactor_adaptive(u + 5 * m, rhoi + NS * m, t_end, h0, rtol, atol, max_steps,
                                    rejects ? rejects + m : nullptr);
}

// Temperature of each reactor (sensible energy inversion)
void orc_reactor_T(long n, const double* u, const double* rhoi, double* T)
{
    for (long m = 0; m < n; ++m) {
        const double* uu = u + 5 * m;
        double rho = uu[0];
        double ux = uu[1] / rho, uy = uu[2] / rho, uz = uu[3] / rho;
        double ei = uu[4] - rho * (0.5 * (ux * ux + uy * uy + uz * uz));
        T[m] = T_given_ey(ei, rhoi + NS * m, nullptr);
    }
}

} // extern "C"
