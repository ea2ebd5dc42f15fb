// Test infrastructure only (never linked into the product).
//
// Our own driver around the reference QWalk objects, compiled together with the reference
// sources by oracle/Makefile.ref.  It writes golden vectors: for each fixed electron
// configuration it builds a fresh Wavefunction and Sample_point (so no incremental state
// carries over) and prints, with %.17g,
//   ln|Psi|, sign, per electron grad(Psi)/Psi (3) and lap(Psi)/Psi,
//   kinetic energy, local (Coulomb) potential, nonlocal pseudopotential energy,
//   the linear-parameter derivatives of Wavefunction::getParmDeriv (gradient, gradderiv,
//   val_gradient) with the parameter values of Wavefunction_data::getVarParms,
//   and the T-move candidates of Pseudopotential::calcNonlocTmove (electron, displacement,
//   vxx) with the nonlocal energy it returns (the terms with vxx >= 0 only).
// The pseudopotential is made deterministic with the identity quadrature rotation, as the
// reference allows (src/system/Pseudopotential.h:80 setDeterministic,
// src/system/Pseudopotential.cpp:828 rotateQuadrature).
//
// Usage: qw_golden <deck> <configs>
//   deck    : text with SYSTEM{...}, TRIALFUNC{...} and optional PSEUDO{...} sections
//   configs : "K Ne" then K*Ne lines "x y z"
// Output   : one block per configuration, see print_config() below.

#include "Qmc_std.h"
#include "qmc_io.h"
#include "System.h"
#include "Sample_point.h"
#include "Wavefunction.h"
#include "Wavefunction_data.h"
#include "Pseudopotential.h"
#include <cstdio>
#include <fstream>

static void print_config(int k, int ne, Wf_return& val, Array2<doublevar>& lap, doublevar kin,
                         doublevar vloc, doublevar vnl) {
  printf("config %d\n", k);
  printf("logpsi %.17g\n", val.amp(0, 0));
  printf("sign %d\n", val.sign(0));
  for (int e = 0; e < ne; e++)
    printf("elec %d %.17g %.17g %.17g %.17g\n", e, lap(e, 0), lap(e, 1), lap(e, 2), lap(e, 3));
  printf("kinetic %.17g\n", kin);
  printf("vloc %.17g\n", vloc);
  printf("vnonloc %.17g\n", vnl);
}

int main(int argc, char** argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: qw_golden deck configs\n");
    return 2;
  }
  std::ifstream deck(argv[1]);
  if (!deck) {
    fprintf(stderr, "cannot open %s\n", argv[1]);
    return 2;
  }
  vector<string> words;
  parsefile(deck, words);

  vector<vector<string> > pseudotext;
  vector<string> sec;
  unsigned int pos = 0;
  while (readsection(words, pos, sec, "PSEUDO") != 0) pseudotext.push_back(sec);
  vector<string> systext, wftext;
  pos = 0;
  if (!readsection(words, pos, systext, "SYSTEM")) {
    fprintf(stderr, "no SYSTEM section\n");
    return 2;
  }
  pos = 0;
  if (!readsection(words, pos, wftext, "TRIALFUNC")) {
    fprintf(stderr, "no TRIALFUNC section\n");
    return 2;
  }

  System* sys = NULL;
  allocate(systext, sys);
  Pseudopotential* psp = NULL;
  sys->generatePseudo(pseudotext, psp);
  psp->setDeterministic(1);
  Array1<doublevar> x(3), y(3), z(3);
  x = 0.0; y = 0.0; z = 0.0;
  x(0) = 1.0; y(1) = 1.0; z(2) = 1.0;
  psp->rotateQuadrature(x, y, z);
  Wavefunction_data* wfdata = NULL;
  allocate(wftext, sys, wfdata);

  std::ifstream cfg(argv[2]);
  int K = 0, ne = 0;
  cfg >> K >> ne;
  if (ne != sys->nelectrons(0) + sys->nelectrons(1)) {
    fprintf(stderr, "electron count mismatch %d vs %d\n", ne,
            sys->nelectrons(0) + sys->nelectrons(1));
    return 2;
  }
  for (int k = 0; k < K; k++) {
    Array2<doublevar> epos(ne, 3);
    for (int e = 0; e < ne; e++)
      for (int d = 0; d < 3; d++) cfg >> epos(e, d);
    Wavefunction* wf = NULL;
    wfdata->generateWavefunction(wf);
    Sample_point* sample = NULL;
    sys->generateSample(sample);
    sample->attachObserver(wf);
    Array1<doublevar> p(3);
    for (int e = 0; e < ne; e++) {
      for (int d = 0; d < 3; d++) p(d) = epos(e, d);
      sample->setElectronPos(e, p);
    }
    wf->notify(all_electrons_move, 0);
    wf->updateLap(wfdata, sample);
    int nwf = wf->nfunc();
    Wf_return val(nwf, 5);
    wf->getVal(wfdata, 0, val);
    Array2<doublevar> lap(ne, 4);
    Wf_return tmp(nwf, 5);
    for (int e = 0; e < ne; e++) {
      wf->getLap(wfdata, e, tmp);
      for (int d = 0; d < 4; d++) lap(e, d) = tmp.amp(0, d + 1);
    }
    Array1<doublevar> kin(nwf);
    sys->calcKinetic(wfdata, sample, wf, kin);
    doublevar vloc = sys->calcLoc(sample);
    Array1<doublevar> test(psp->nTest());
    test = 0.0;
    Array1<doublevar> vnl(nwf);
    vnl = 0.0;
    psp->calcNonlocWithTest(wfdata, sys, sample, wf, test, vnl);
    print_config(k, ne, val, lap, kin(0), vloc, vnl(0));
    // T-move candidates (src/system/Pseudopotential.cpp:371-389,752-763)
    vector<Tmove> tmov;
    Array1<doublevar> vnl_t(nwf);
    psp->calcNonlocTmove(wfdata, sys, sample, wf, vnl_t, tmov);
    printf("vnl_tmove %.17g\n", vnl_t(0));
    for (size_t t = 0; t < tmov.size(); t++)
      printf("tmove %d %.17g %.17g %.17g %.17g\n", tmov[t].e, tmov[t].vxx, tmov[t].pos(0), tmov[t].pos(1),
             tmov[t].pos(2));
    // linear-parameter derivatives (Wavefunction::getParmDeriv) and the parameter values
    Parm_deriv_return pd;
    pd.need_hessian = 0;
    const int pdok = wf->getParmDeriv(wfdata, sample, pd);
    Array1<doublevar> parms;
    wfdata->getVarParms(parms);
    const int np = pd.gradient.GetDim(0);
    printf("pdok %d\n", pdok);
    printf("nparm %d\n", np);
    for (int q = 0; q < parms.GetDim(0); q++) printf("parm %d %.17g\n", q, parms(q));
    for (int q = 0; q < np; q++) printf("pdgrad %d %.17g\n", q, pd.gradient(q));
    for (int q = 0; q < np; q++)
      for (int e = 0; e < ne; e++)
        printf("pdgd %d %d %.17g %.17g %.17g %.17g\n", q, e, pd.gradderiv(q, e, 0), pd.gradderiv(q, e, 1),
               pd.gradderiv(q, e, 2), pd.gradderiv(q, e, 3));
    for (int e = 0; e < ne && pd.val_gradient.GetDim(0) == ne; e++)
      printf("pdvg %d %.17g %.17g %.17g\n", e, pd.val_gradient(e, 0), pd.val_gradient(e, 1), pd.val_gradient(e, 2));
    delete wf;
    delete sample;
  }
  delete psp;
  delete wfdata;
  delete sys;
  return 0;
}
