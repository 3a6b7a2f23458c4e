// C++ host program over the C ABI (built with g++ only: no HIP headers),
// exercising the reference-style classes of mgc.hpp the way the reference's
// drivers do: Aggregation::Create, CoarsenedMatrix::CoarsenOperator, the
// runChecks of include/MultigridPreconditioner.h:311-446, and an FGMRES solve
// preconditioned by the multigrid hierarchy (src/Test_wilsonclover_mg.cc).
// Prints one line per check and exits non-zero on the first failure.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../grid-mg-bench_amd/include/mgc/mgc.hpp"

using cd = std::complex<double>;

static std::vector<cd> host(const mgc::DeviceField& f) {
  std::vector<cd> h(f.count());
  f.download(h.data());
  return h;
}

static double rel(const std::vector<cd>& a, const std::vector<cd>& b) {
  double n = 0, d = 0;
  for (size_t i = 0; i < a.size(); ++i) {
    n += std::norm(a[i] - b[i]);
    d += std::norm(a[i]);
  }
  return std::sqrt(n / d);
}

static int failures = 0;
static void expect(const char* what, double value, double bound) {
  const bool ok = value <= bound;
  std::printf("%-48s %.3e %s %.1e %s\n", what, value, ok ? "<=" : ">", bound, ok ? "PASS" : "FAIL");
  std::fflush(stdout);
  if (!ok) ++failures;
}

int main() {
  try {
    const uint64_t seed = 1ull | (2ull << 16) | (3ull << 32) | (4ull << 48);
    const mgc::Coordinate fine{{8, 8, 8, 8}}, block{{4, 4, 4, 4}};
    const int nb = 12, N = 2 * nb;
    const int Vf = fine.volume(), Vc = Vf / block.volume();
    const mgc::Coordinate coarse{{2, 2, 2, 2}};

    // fine operator: Wilson-clover m = -0.1, csw = 1.0 on a random SU(3) field
    mgc::DeviceField U((size_t)4 * Vf * 9, MGC_F64);
    mgc::check(mgc_random_gauge(U.data(), fine.v, 0.4, seed, 1, nullptr));
    mgc::WilsonCloverOperator D(fine, -0.1, 1.0);
    D.ImportGauge(U);

    // M = Mdiag + sum_mu (Mdir_+mu + Mdir_-mu)
    mgc::DeviceField v((size_t)Vf * 12, MGC_F64), w((size_t)Vf * 12, MGC_F64);
    v.gaussian(seed, 2);
    D.M(v, w);
    std::vector<cd> Mv = host(w), sum(Vf * 12, 0.0);
    D.OpDiag(v, w);
    {
      std::vector<cd> x = host(w);
      for (size_t i = 0; i < x.size(); ++i) sum[i] += x[i];
    }
    for (int mu = 0; mu < 4; ++mu)
      for (int disp : {+1, -1}) {
        D.OpDir(v, w, mu, disp);
        std::vector<cd> x = host(w);
        for (size_t i = 0; i < x.size(); ++i) sum[i] += x[i];
      }
    expect("fine: M - (Mdiag + sum Mdir)", rel(Mv, sum), 1e-14);

    // Aggregation::Create + CoarsenOperator
    mgc::DeviceField basis((size_t)nb * Vf * 12, MGC_F64);
    basis.gaussian(seed, 3);
    mgc::Aggregation agg(fine, block, nb);
    agg.Create(basis, 2);
    std::vector<double> orth = agg.CheckOrthogonal();
    double worst = 0;
    for (double e : orth) worst = std::fmax(worst, e);
    expect("aggregation: CheckOrthogonal max norm2", worst, 1e-24);
    mgc::CoarsenedMatrix Dc(coarse, N, MGC_F64);
    Dc.CoarsenOperator(D, agg, MGC_COARSEN_FULL);

    // (1 - R P) v_c = 0 and R D P v_c = D_c v_c
    mgc::DeviceField vc((size_t)Vc * N, MGC_F64), wc((size_t)Vc * N, MGC_F64),
        uc((size_t)Vc * N, MGC_F64);
    vc.gaussian(seed, 4);
    agg.PromoteFromSubspace(vc, v);
    agg.ProjectToSubspace(wc, v);
    expect("transfer: (1 - R P) v_c", rel(host(vc), host(wc)), 1e-13);
    D.M(v, w);
    agg.ProjectToSubspace(wc, w);
    Dc.M(vc, uc);
    expect("Galerkin: R D P v_c - D_c v_c", rel(host(wc), host(uc)), 1e-12);
    // Mdiag + Dhop = M and MooeeInv Mooee = 1 on the coarse operator
    Dc.Mdiag(vc, wc);
    Dc.Dhop(vc, uc, 0);
    std::vector<cd> a = host(wc), b = host(uc);
    for (size_t i = 0; i < a.size(); ++i) a[i] += b[i];
    Dc.M(vc, uc);
    expect("coarse: M - (Mdiag + Dhop)", rel(host(uc), a), 1e-14);
    Dc.Mooee(vc, wc);
    Dc.MooeeInv(wc, uc);
    expect("coarse: MooeeInv Mooee - 1", rel(host(vc), host(uc)), 1e-11);

    // coarse solves in a batch: GMRES to 1e-5, and 24 fixed steps as a HIP graph
    {
      const int nrhs = 2;
      void* cs = nullptr;
      mgc::check(mgc_stream_create(&cs));
      {
        mgc::DeviceField bc((size_t)Vc * nrhs * N, MGC_F32), x1((size_t)Vc * nrhs * N, MGC_F32),
            x2((size_t)Vc * nrhs * N, MGC_F32);
        bc.gaussian(seed, 6, cs);
        mgc::CoarseSolveBatch tolb({&Dc}, nrhs, MGC_F32, 20, cs);
        mgc::CoarseSolveBatch fixb({&Dc}, nrhs, MGC_F32, 20, cs, 24, 2);
        const std::vector<mgc_solve_stats> s1 = tolb.Solve({&bc}, {&x1}, 1e-5);
        const std::vector<mgc_solve_stats> s2 = fixb.Solve({&bc}, {&x2}, 1e-5);
        double r1 = 0, r2 = 0;
        for (int r = 0; r < nrhs; ++r) {
          r1 = std::fmax(r1, s1[r].true_residual);
          r2 = std::fmax(r2, s2[r].true_residual);
        }
        std::printf("coarse batch: GMRES %d iterations residual %.2e; 24 fixed steps residual "
                    "%.2e\n", s1[0].iterations, r1, r2);
        expect("coarse batch: GMRES residual", r1, 2e-5);
        expect("coarse batch: fixed-step graph residual", r2, 1e-4);
        mgc::check(mgc_stream_synchronize(cs));
      }  // the batches and fields go before their stream
      mgc::check(mgc_stream_destroy(cs));
    }

    // two-level hierarchy and a solve to 1e-10
    mgc_mg_params p;
    mgc::check(mgc_mg_params_default(&p));
    p.nlevels = 2;
    p.nb[0] = 8;
    mgc::MultiGridPreconditioner mg(D, p);
    mg.initialSetup();
    mgc::DeviceField src((size_t)Vf * 12, MGC_F64), sol((size_t)Vf * 12, MGC_F64);
    src.gaussian(seed, 5);
    const mgc_solve_stats st = mg.Solve(src, sol, 1e-10, 500, 20);
    D.M(sol, w);
    std::printf("solve: %d outer iterations, %.3f s, reported residual %.3e\n", st.iterations,
                st.seconds, st.true_residual);
    expect("solve: ||b - D x|| / ||b||", rel(host(src), host(w)), 1.2e-10);
  } catch (const mgc::Exception& e) {
    std::printf("mgc::Exception %d: %s\n", e.code, e.what());
    return 2;
  }
  std::printf("%s\n", failures ? "FAILED" : "ALL PASSED");
  return failures ? 1 : 0;
}
