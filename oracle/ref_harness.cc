// Golden-vector harness: drives the reference OSCARS SR C++ engine (compiled
// from its own sources under /root/reference/src by oracle/Makefile, target
// `ref`) and writes raw float64 results that tests/golden/make_golden.py turns
// into committed fixtures.  Test infrastructure only: it is never shipped and
// the product never links it.
//
// Usage:  ref_harness <mode> key=value ...
//   modes: traj | spectrum | flux | pd | totalpower | beam | multiflux
//   see parse() below for the keys.  Output (key out=PATH) is a flat float64
//   little-endian file; its layout per mode is documented in write_*().
//
// The calls mirror what the Python entry points do:
//   set_particle_beam        OSCARSSR_Python.cc:1641-1655 (+ sigma :1822-1829)
//   add_bfield_undulator     OSCARSSR_Python.cc:736-818
//   add_bfield_gaussian      OSCARSSR_Python.cc:561-636
//   add_bfield_uniform       OSCARSSR_Python.cc:640-718
//   calculate_spectrum       OSCARSSR_Python.cc:2215-2351
//   calculate_flux_rectangle OSCARSSR_Python.cc:3009-3233
//   calculate_power_density_rectangle OSCARSSR_Python.cc:2587-2807
//   calculate_total_power    OSCARSSR_Python.cc:2369-2398

#include "OSCARSSR.h"
#include "TField.h"
#include "TField3D_Gaussian.h"
#include "TField3D_IdealUndulator.h"
#include "TField3D_UniformBox.h"
#include "TSpectrumContainer.h"
#include "T3DScalarContainer.h"
#include "TSurfacePoints_Rectangle.h"
#include "TSurfacePoints_3D.h"
#include "TVector2D.h"
#include "TRandomA.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <map>
#include <sstream>
#include <string>
#include <vector>

extern TRandomA* gRandomA;

// Ideal planar undulator whose 2*NPeriods interior half-period poles are each
// scaled by (1 + delta_p).  Terminations (0.25 / 0.75 half periods) are not
// scaled.  Same envelope and sine as TField3D_IdealUndulator::GetF
// (TField3D_IdealUndulator.cc:106-139); the pole index is
// p = floor((D - (-L/2 + PS + lambda)) / (lambda/2)).
class PoleScaledUndulator : public TField {
 public:
  PoleScaledUndulator(TVector3D const& F, TVector3D const& P, int N, std::vector<double> const& S)
      : fF(F), fP(P), fN(N), fScale(S) {
    fLam = fP.Mag();
    fU = fP.UnitVector();
    fLen = fP.Mag() * (fN + 2);
  }
  double GetFx(double x, double y, double z) const { return GetF(TVector3D(x, y, z)).GetX(); }
  double GetFy(double x, double y, double z) const { return GetF(TVector3D(x, y, z)).GetY(); }
  double GetFz(double x, double y, double z) const { return GetF(TVector3D(x, y, z)).GetZ(); }
  TVector3D GetF(double x, double y, double z) const { return GetF(TVector3D(x, y, z)); }
  TVector3D GetF(TVector3D const& X) const {
    double const D = X.Dot(fU);
    double const s = sin(2.0 * 3.14159265358979323846 * D / fLam);
    TVector3D F(0, 0, 0);
    if (D > fLen / 2. || D < -fLen / 2.) return F;
    if (D < -fLen / 2. + fLam || D > fLen / 2. - fLam) {
      if (D < -fLen / 2. + fLam / 2. || D > fLen / 2. - fLam / 2.) {
        return TVector3D((0.25 * fF.GetX()) * s, (0.25 * fF.GetY()) * s, (0.25 * fF.GetZ()) * s);
      }
      return TVector3D((0.75 * fF.GetX()) * s, (0.75 * fF.GetY()) * s, (0.75 * fF.GetZ()) * s);
    }
    int p = (int)floor((D - (-fLen / 2. + fLam)) / (fLam / 2.));
    if (p < 0) p = 0;
    if (p > 2 * fN - 1) p = 2 * fN - 1;
    double const sc = 1.0 + fScale[p];
    return TVector3D((fF.GetX() * s) * sc, (fF.GetY() * s) * sc, (fF.GetZ() * s) * sc);
  }

 private:
  TVector3D fF, fP, fU;
  int fN;
  double fLam, fLen;
  std::vector<double> fScale;
};

static std::map<std::string, std::string> gArgs;

static std::vector<double> vec(std::string const& k, std::vector<double> def = std::vector<double>()) {
  if (!gArgs.count(k)) return def;
  std::vector<double> v;
  std::stringstream ss(gArgs[k]);
  std::string tok;
  while (std::getline(ss, tok, ',')) v.push_back(strtod(tok.c_str(), 0));
  return v;
}
static double num(std::string const& k, double def) { return gArgs.count(k) ? strtod(gArgs[k].c_str(), 0) : def; }
static std::string str(std::string const& k, std::string const& def) { return gArgs.count(k) ? gArgs[k] : def; }
static TVector3D v3(std::string const& k, TVector3D def) {
  std::vector<double> v = vec(k);
  return v.size() == 3 ? TVector3D(v[0], v[1], v[2]) : def;
}

// Zero a stretch of the stack below the caller's frame.  CalculateTrajectory stores its
// uninitialised dxdt as the first point's A/c (OSCARSSR.cc:604-630); called right after
// this, that array sits in zeroed memory, so the value is 0 (the restatement's convention)
// instead of leftover bytes.
__attribute__((noinline)) static void scrub_stack() {
  volatile char buf[1 << 16];
  for (size_t i = 0; i < sizeof(buf); ++i) buf[i] = 0;
}

static FILE* gOut = 0;
static void put(double v) { fwrite(&v, sizeof(double), 1, gOut); }

static void setup(OSCARSSR& osr) {
  // Undulators: und1=Bx,By,Bz,Px,Py,Pz,N[,phase[,taper[,cx,cy,cz]]], und2=...
  for (int k = 1; k <= 4; ++k) {
    std::vector<double> u = vec("und" + std::to_string(k));
    if (u.size() < 7) continue;
    double phase = u.size() > 7 ? u[7] : 0;
    double taper = u.size() > 8 ? u[8] : 0;
    TVector3D ctr = u.size() > 11 ? TVector3D(u[9], u[10], u[11]) : TVector3D(0, 0, 0);
    osr.AddMagneticField((TField*)new TField3D_IdealUndulator(TVector3D(u[0], u[1], u[2]), TVector3D(u[3], u[4], u[5]), (int)u[6], ctr, phase, taper));
  }
  // Gaussian fields: gauss1=Bx,By,Bz,sx,sy,sz[,cx,cy,cz[,rx,ry,rz]] (then gauss2), and
  // uniform boxes: box1=Bx,By,Bz[,wx,wy,wz[,cx,cy,cz[,rx,ry,rz]]] (then box2), added
  // after the undulators in this order; the field sum runs in insertion order.
  for (int k = 1; k <= 2; ++k) {
    std::vector<double> g = vec("gauss" + std::to_string(k));
    if (g.size() < 6) continue;
    TVector3D ctr = g.size() >= 9 ? TVector3D(g[6], g[7], g[8]) : TVector3D(0, 0, 0);
    TVector3D rot = g.size() >= 12 ? TVector3D(g[9], g[10], g[11]) : TVector3D(0, 0, 0);
    osr.AddMagneticField((TField*)new TField3D_Gaussian(TVector3D(g[0], g[1], g[2]), ctr, TVector3D(g[3], g[4], g[5]), rot));
  }
  for (int k = 1; k <= 2; ++k) {
    std::vector<double> b = vec("box" + std::to_string(k));
    if (b.size() < 3) continue;
    TVector3D w = b.size() >= 6 ? TVector3D(b[3], b[4], b[5]) : TVector3D(0, 0, 0);
    TVector3D ctr = b.size() >= 9 ? TVector3D(b[6], b[7], b[8]) : TVector3D(0, 0, 0);
    TVector3D rot = b.size() >= 12 ? TVector3D(b[9], b[10], b[11]) : TVector3D(0, 0, 0);
    osr.AddMagneticField((TField*)new TField3D_UniformBox(TVector3D(b[0], b[1], b[2]), w, ctr, rot));
  }
  // Electric fields (OSCARSSR::AddElectricField, as add_efield_undulator / _gaussian /
  // _uniform add them, OSCARSSR_Python.cc:994-1240): eund1=Ex,Ey,Ez,Px,Py,Pz,N[,phase[,cx,cy,cz]]
  // (no taper), egauss1= and ebox1= as gauss1= and box1=; added in this order.
  for (int k = 1; k <= 2; ++k) {
    std::vector<double> u = vec("eund" + std::to_string(k));
    if (u.size() < 7) continue;
    TVector3D ctr = u.size() > 10 ? TVector3D(u[8], u[9], u[10]) : TVector3D(0, 0, 0);
    osr.AddElectricField((TField*)new TField3D_IdealUndulator(TVector3D(u[0], u[1], u[2]), TVector3D(u[3], u[4], u[5]),
                                                              (int)u[6], ctr, u.size() > 7 ? u[7] : 0));
  }
  for (int k = 1; k <= 2; ++k) {
    std::vector<double> g = vec("egauss" + std::to_string(k));
    if (g.size() < 6) continue;
    TVector3D ctr = g.size() >= 9 ? TVector3D(g[6], g[7], g[8]) : TVector3D(0, 0, 0);
    TVector3D rot = g.size() >= 12 ? TVector3D(g[9], g[10], g[11]) : TVector3D(0, 0, 0);
    osr.AddElectricField((TField*)new TField3D_Gaussian(TVector3D(g[0], g[1], g[2]), ctr, TVector3D(g[3], g[4], g[5]), rot));
  }
  for (int k = 1; k <= 2; ++k) {
    std::vector<double> b = vec("ebox" + std::to_string(k));
    if (b.size() < 3) continue;
    TVector3D w = b.size() >= 6 ? TVector3D(b[3], b[4], b[5]) : TVector3D(0, 0, 0);
    TVector3D ctr = b.size() >= 9 ? TVector3D(b[6], b[7], b[8]) : TVector3D(0, 0, 0);
    TVector3D rot = b.size() >= 12 ? TVector3D(b[9], b[10], b[11]) : TVector3D(0, 0, 0);
    osr.AddElectricField((TField*)new TField3D_UniformBox(TVector3D(b[0], b[1], b[2]), w, ctr, rot));
  }
  // Field maps read by the reference's TField3D_Grid (OSCARSSR::AddMagneticField,
  // OSCARSSR.cc:68-90): gridfile1=path gridfmt1=format [gridrot1=rx,ry,rz]
  // [gridtrans1=tx,ty,tz] [gridscale1=s0,s1,...] (then gridfile2 ...), and one map
  // interpolated in a parameter (OSCARSSR::AddMagneticFieldInterpolated, :94-114):
  // gridmap=p0:path0;p1:path1;... gridmapfmt=format gridparam=p [gridmaprot=...]
  // [gridmaptrans=...] [gridmapscale=...]; added after the boxes, in this order.
  for (int k = 1; k <= 2; ++k) {
    std::string const key = "gridfile" + std::to_string(k);
    if (!gArgs.count(key)) continue;
    std::string const sk = std::to_string(k);
    osr.AddMagneticField(str(key, ""), str("gridfmt" + sk, "OSCARS"), v3("gridrot" + sk, TVector3D(0, 0, 0)),
                         v3("gridtrans" + sk, TVector3D(0, 0, 0)), vec("gridscale" + sk));
  }
  if (gArgs.count("gridmap")) {
    std::vector<std::pair<double, std::string> > mapping;
    std::string m = str("gridmap", "");
    size_t p0 = 0;
    while (p0 < m.size()) {
      size_t const p1 = m.find(';', p0);
      std::string const e = m.substr(p0, p1 == std::string::npos ? std::string::npos : p1 - p0);
      size_t const c = e.find(':');
      mapping.push_back(std::make_pair(atof(e.substr(0, c).c_str()), e.substr(c + 1)));
      if (p1 == std::string::npos) break;
      p0 = p1 + 1;
    }
    osr.AddMagneticFieldInterpolated(mapping, str("gridmapfmt", "OSCARS"), num("gridparam", 0),
                                     v3("gridmaprot", TVector3D(0, 0, 0)), v3("gridmaptrans", TVector3D(0, 0, 0)),
                                     vec("gridmapscale"));
  }
  // Pole-scaled undulator: poles=Bx,By,Bz,Px,Py,Pz,N  scales=d0,d1,...
  std::vector<double> pu = vec("poles");
  if (pu.size() == 7) {
    osr.AddMagneticField((TField*)new PoleScaledUndulator(TVector3D(pu[0], pu[1], pu[2]), TVector3D(pu[3], pu[4], pu[5]), (int)pu[6], vec("scales")));
  }
  // Beam, as set_particle_beam does.
  TVector3D x0 = v3("x0", TVector3D(0, 0, 0));
  TVector3D d0 = v3("d0", TVector3D(0, 0, 1));
  osr.ClearParticleBeams();
  osr.AddParticleBeam(str("type", "electron"), "beam", x0, d0, num("energy", 3), num("t0", 0), num("current", 0.5), 1);
  std::vector<double> beta = vec("beta"), emit = vec("emit");
  TVector2D su(0, 0), sup(0, 0);
  if (beta.size() == 2 && emit.size() == 2) {
    su[0] = sqrt(emit[0] * beta[0]);
    su[1] = sqrt(emit[1] * beta[1]);
    sup[0] = sqrt(emit[0] / beta[0]);
    sup[1] = sqrt(emit[1] / beta[1]);
  }
  TVector3D hdir = gArgs.count("hdir") ? v3("hdir", TVector3D(1, 0, 0)) : d0.Orthogonal().UnitVector();
  hdir = hdir.UnitVector();
  osr.GetParticleBeam("beam").SetSigma(hdir, su, sup, v3("latref", TVector3D(0, 0, 0)), num("sigE", 0));
  // A second beam for the weighted pick (TParticleBeamContainer::GetRandomBeamIndexByWeight):
  // beam2=x0x,x0y,x0z,d0x,d0y,d0z,energy,current weight1=w weight2=w; the same sigma
  // settings as the first beam (add_particle_beam, OSCARSSR_Python.cc:1661-1870)
  std::vector<double> b2 = vec("beam2");
  if (b2.size() == 8) {
    osr.ClearParticleBeams();
    osr.AddParticleBeam(str("type", "electron"), "beam", x0, d0, num("energy", 3), num("t0", 0), num("current", 0.5),
                        num("weight1", 1));
    osr.GetParticleBeam("beam").SetSigma(hdir, su, sup, v3("latref", TVector3D(0, 0, 0)), num("sigE", 0));
    TVector3D const d2(b2[3], b2[4], b2[5]);
    osr.AddParticleBeam(str("type", "electron"), "beam2", TVector3D(b2[0], b2[1], b2[2]), d2, b2[6], num("t0", 0),
                        b2[7], num("weight2", 1));
    TVector3D const h2 = gArgs.count("hdir") ? hdir : d2.Orthogonal().UnitVector();
    osr.GetParticleBeam("beam2").SetSigma(h2, su, sup, v3("latref", TVector3D(0, 0, 0)), num("sigE", 0));
  }
  osr.SetNewParticle("beam", "ideal");
  std::vector<double> ct = vec("ct", std::vector<double>{0, 3});
  osr.SetCTStartStop(ct[0], ct[1]);
  if (gArgs.count("npoints")) osr.SetNPointsTrajectory((size_t)num("npoints", 0));
  if (gArgs.count("nthreads_global")) osr.SetNThreadsGlobal((int)num("nthreads_global", 2));
  if (gArgs.count("seed")) osr.SetSeed((int)num("seed", 0));
}

static TSurfacePoints_Rectangle rect() {
  TSurfacePoints_Rectangle S;
  std::vector<double> n = vec("nx", std::vector<double>{11, 11});
  std::vector<double> w = vec("width", std::vector<double>{0.01, 0.01});
  std::vector<double> c = vec("x0x1x2");
  if (c.size() == 9) {
    // three corners, rotated then translated as calculate_*_rectangle does with x0x1x2
    // (OSCARSSR_Python.cc:3128-3151), then TSurfacePoints_Rectangle::Init(NX1, NX2, X0, X1, X2, Normal)
    TVector3D const rot = v3("rot", TVector3D(0, 0, 0)), tr = v3("trans", TVector3D(0, 0, 0));
    std::vector<TVector3D> X;
    for (int k = 0; k < 3; ++k) {
      X.push_back(TVector3D(c[3 * k], c[3 * k + 1], c[3 * k + 2]));
      X.back().RotateSelfXYZ(rot);
      X.back() += tr;
    }
    S.Init((int)n[0], (int)n[1], X[0], X[1], X[2], (int)num("normal", 0));
    return S;
  }
  S.Init(str("plane", "XY"), (int)n[0], (int)n[1], w[0], w[1], v3("rot", TVector3D(0, 0, 0)), v3("trans", TVector3D(0, 0, 30)), (int)num("normal", 0));
  return S;
}

// Point list as calculate_flux / calculate_power_density build it (OSCARSSR_Python.cc:
// 2472-2506, 2911-2940): points=x,y,z,nx,ny,nz,...; each point and normal rotated by
// prot (RotateSelfXYZ), the normal reversed when flip is set and pnormal == -1, the
// point translated by ptrans.
static TSurfacePoints_3D points3d(bool flip) {
  std::vector<double> p = vec("points");
  TVector3D const rot = v3("prot", TVector3D(0, 0, 0)), tr = v3("ptrans", TVector3D(0, 0, 0));
  int const nd = (int)num("pnormal", 0);
  TSurfacePoints_3D S;
  for (size_t i = 0; i + 5 < p.size(); i += 6) {
    TVector3D X(p[i], p[i + 1], p[i + 2]), N(p[i + 3], p[i + 4], p[i + 5]);
    X.RotateSelfXYZ(rot);
    N.RotateSelfXYZ(rot);
    if (flip && nd == -1) N *= -1;
    X += tr;
    S.AddPoint(X, N);
  }
  return S;
}

static void write_container(T3DScalarContainer const& C) {
  // layout: N, then N x (x, y, z, v)
  put((double)C.GetNPoints());
  for (size_t i = 0; i != C.GetNPoints(); ++i) {
    put(C.GetPoint(i).GetX().GetX());
    put(C.GetPoint(i).GetX().GetY());
    put(C.GetPoint(i).GetX().GetZ());
    put(C.GetPoint(i).GetV());
  }
}

int main(int argc, char** argv) {
  if (argc < 2) {
    fprintf(stderr, "usage: ref_harness mode key=value...\n");
    return 2;
  }
  std::string mode = argv[1];
  for (int i = 2; i < argc; ++i) {
    std::string a = argv[i];
    size_t e = a.find('=');
    if (e != std::string::npos) gArgs[a.substr(0, e)] = a.substr(e + 1);
  }
  gOut = fopen(str("out", "/dev/stdout").c_str(), "wb");
  if (!gOut) return 3;

  OSCARSSR osr;
  setup(osr);
  int const nthreads = (int)num("nthreads", 1);
  int const nparticles = (int)num("nparticles", 0);
  // gpu=1: the reference's GPU routes (a harness built with -DCUDA and an OSCARSSR_Cuda_*
  // implementation; the plain build throws "GPU functionality not compiled ...")
  int const gpu = (int)num("gpu", 0);

  try {
    if (mode == "traj") {
      // layout: n, dt, then x[n] y[n] z[n] bx by bz ax ay az
      osr.CalculateTrajectory();
      TParticleTrajectoryPoints const& T = osr.GetTrajectory();
      size_t n = T.GetNPoints();
      put((double)n);
      put(T.GetDeltaT());
      for (int c = 0; c < 9; ++c)
        for (size_t i = 0; i < n; ++i) {
          TVector3D const& V = c < 3 ? T.GetX(i) : (c < 6 ? T.GetB(i) : T.GetAoverC(i));
          put(V[c % 3]);
        }
    } else if (mode == "spectrum") {
      // layout: NE, then NE x (E, flux)
      TSpectrumContainer S;
      std::vector<double> pts = vec("points_ev");
      if (pts.size()) {
        S.Init(pts);
      } else {
        std::vector<double> er = vec("erange", std::vector<double>{100, 3000});
        S.Init((size_t)num("ne", 100), er[0], er[1]);
      }
      osr.CalculateSpectrum(v3("obs", TVector3D(0, 0, 30)), S, nparticles, nthreads, gpu);
      // file output as calculate_spectrum(ofile=...) writes it (OSCARSSR_Python.cc:2340-2343)
      if (gArgs.count("ofile")) S.WriteToFileText(str("ofile", ""));
      put((double)S.GetNPoints());
      for (size_t i = 0; i < S.GetNPoints(); ++i) {
        put(S.GetEnergy(i));
        put(S.GetFlux(i));
      }
    } else if (mode == "flux") {
      T3DScalarContainer C;
      if (gArgs.count("points")) {
        // points=x,y,z,nx,ny,nz,x,y,z,... (calculate_flux with a point list, dim=3)
        TSurfacePoints_3D S = points3d(false);
        osr.CalculateFlux(S, num("energy_ev", 100), C, nparticles, nthreads, gpu, 3, "");
      } else {
        TSurfacePoints_Rectangle S = rect();
        osr.CalculateFlux(S, num("energy_ev", 100), C, nparticles, nthreads, gpu, (int)num("dim", 2), str("ofile", ""));
      }
      write_container(C);
    } else if (mode == "pd") {
      T3DScalarContainer C;
      bool const directional = num("normal", 0) != 0;
      if (gArgs.count("points")) {
        TSurfacePoints_3D S = points3d(true);
        osr.CalculatePowerDensity(S, C, 3, num("pnormal", 0) != 0, nparticles, "", nthreads, gpu);
      } else {
        TSurfacePoints_Rectangle S = rect();
        osr.CalculatePowerDensity(S, C, (int)num("dim", 2), directional, nparticles, str("ofile", ""), nthreads, gpu);
      }
      write_container(C);
    } else if (mode == "flux1") {
      // OSCARSSR::CalculateFlux1 (no Python binding; the caller adds the container's
      // points, 2-D as its hard-coded Dimension): the container, then the trajectory
      // it computed and used (layout of "traj"), whose first A/c is the reference's
      // uninitialised value
      TSurfacePoints_Rectangle S = rect();
      T3DScalarContainer C;
      for (size_t i = 0; i != S.GetNPoints(); ++i) C.AddPoint(TVector3D(S.GetX1(i), S.GetX2(i), 0), 0);
      TParticleA P = osr.GetCurrentParticle();
      if (num("scrub", 1) != 0) scrub_stack();
      osr.CalculateFlux1(P, S, num("energy_ev", 100), C, str("ofile", ""));
      write_container(C);
      TParticleTrajectoryPoints& T = P.GetTrajectory();
      size_t const n = T.GetNPoints();
      put((double)n);
      put(T.GetDeltaT());
      for (int c = 0; c < 9; ++c)
        for (size_t i = 0; i < n; ++i) {
          TVector3D const& V = c < 3 ? T.GetX(i) : (c < 6 ? T.GetB(i) : T.GetAoverC(i));
          put(V[c % 3]);
        }
    } else if (mode == "totalpower") {
      put(osr.CalculateTotalPower());
    } else if (mode == "efield") {
      // calculate_efield_vs_time (OSCARSSR_Python.cc:3573-3630): container points = E,
      // values = arrival time
      T3DScalarContainer XYZT;
      osr.CalculateElectricFieldTimeDomain(v3("obs", TVector3D(0, 0, 30)), XYZT);
      write_container(XYZT);
    } else if (mode == "avg3d") {
      // average_flux / average_power_density (OSCARSSR_Python.cc:3475-3563):
      // ifiles=a,b,...  binary=0|1  dim=2|3  [ofile=, bofile=]
      std::vector<std::string> files;
      std::stringstream ss(str("ifiles", ""));
      std::string tok;
      while (std::getline(ss, tok, ',')) files.push_back(tok);
      int const dim = (int)num("dim", 2);
      T3DScalarContainer C;
      if (num("binary", 0) != 0) C.AverageFromFilesBinary(files, dim);
      else C.AverageFromFilesText(files, dim);
      if (gArgs.count("ofile")) C.WriteToFileText(str("ofile", ""), dim);
      if (gArgs.count("bofile")) C.WriteToFileBinary(str("bofile", ""), dim);
      write_container(C);
    } else if (mode == "avgspec") {
      // average_spectra (OSCARSSR_Python.cc:3252-3318), text files: ifiles=a,b,... [ofile=]
      std::vector<std::string> files;
      std::stringstream ss(str("ifiles", ""));
      std::string tok;
      while (std::getline(ss, tok, ',')) files.push_back(tok);
      TSpectrumContainer S;
      S.AverageFromFilesText(files);
      if (gArgs.count("ofile")) S.WriteToFileText(str("ofile", ""));
      put((double)S.GetNPoints());
      for (size_t i = 0; i < S.GetNPoints(); ++i) {
        put(S.GetEnergy(i));
        put(S.GetFlux(i));
      }
    } else if (mode == "efield_at") {
      // summed electric field at the points of pts=path (OSCARSSR::GetE); layout as "bfield"
      FILE* fp = fopen(str("pts", "").c_str(), "rb");
      if (!fp) return 4;
      std::vector<double> xyz;
      double v;
      while (fread(&v, sizeof(double), 1, fp) == 1) xyz.push_back(v);
      fclose(fp);
      size_t const n = xyz.size() / 3;
      put((double)n);
      for (size_t i = 0; i != n; ++i) {
        TVector3D const E = osr.GetE(TVector3D(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
        put(E.GetX());
        put(E.GetY());
        put(E.GetZ());
      }
    } else if (mode == "bfield") {
      // total field at the points of the raw float64 file pts=path (x, y, z per point):
      // OSCARSSR::GetB (TFieldContainer::GetF); layout: N, then N x (Bx, By, Bz)
      FILE* fp = fopen(str("pts", "").c_str(), "rb");
      if (!fp) return 4;
      std::vector<double> xyz;
      double v;
      while (fread(&v, sizeof(double), 1, fp) == 1) xyz.push_back(v);
      fclose(fp);
      size_t const n = xyz.size() / 3;
      put((double)n);
      for (size_t i = 0; i != n; ++i) {
        TVector3D const B = osr.GetB(TVector3D(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]));
        put(B.GetX());
        put(B.GetY());
        put(B.GetZ());
      }
    } else if (mode == "beam") {
      // layout: count, then count x (x0[3], beta0[3], E0, gamma, q/m/gamma)
      int count = (int)num("count", 8);
      put((double)count);
      for (int i = 0; i < count; ++i) {
        osr.SetNewParticle();
        TParticleA const& P = osr.GetCurrentParticle();
        put(P.GetX0().GetX());
        put(P.GetX0().GetY());
        put(P.GetX0().GetZ());
        put(P.GetB0().GetX());
        put(P.GetB0().GetY());
        put(P.GetB0().GetZ());
        put(P.GetE0());
        put(P.GetGamma());
        put(P.GetQoverMGamma());
      }
    } else if (mode == "uniforms") {
      // layout: count, then count uniforms [0, 1) from the global generator after set_seed
      // (TRandomA::Uniform, TRandomA.cc:61-64)
      int count = (int)num("count", 8);
      put((double)count);
      for (int i = 0; i < count; ++i) put(gRandomA->Uniform());
    } else if (mode == "normals") {
      // layout: count, then count standard normals straight from the global generator after set_seed
      int count = (int)num("count", 8);
      put((double)count);
      for (int i = 0; i < count; ++i) put(osr.GetRandomNormal());
    } else {
      fprintf(stderr, "unknown mode %s\n", mode.c_str());
      return 2;
    }
  } catch (std::exception const& e) {
    fprintf(stderr, "reference threw: %s\n", e.what());
    fclose(gOut);
    return 4;
  }
  fclose(gOut);
  return 0;
}
