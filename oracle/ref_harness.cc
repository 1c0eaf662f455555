// Dump harness around the reference LULESH 2.x sources.
//
// Test infrastructure only.  This translation unit pulls the reference's
// lulesh.cc in where it lies (include path set by oracle/Makefile) so that its
// file-static phase functions can be called one at a time.  The reference's
// own main() is renamed and never called.  Nothing here changes the
// reference's arithmetic: every physics call below is the reference's own
// function, run through its OpenMP gather path (threads > 1,
// lulesh-init.cc:280, lulesh.cc:540-586).
//
// Usage:
//   refdump -s S [-r R] [-b B] [-c C] [-i ITS] [-t THREADS] [-m EMULT]
//           [-d DUMPFILE] [-D c1,c2,...]
// Prints one JSON line with the end-of-run quantities at 17 significant
// digits (the quantities of lulesh-util.cc:175-230) and the loop wall time.
// With -d, writes the set-up arrays and, for every cycle listed in -D, the
// state after LagrangeNodal, after the monotonic-Q gradients, after
// LagrangeElements and after CalcTimeConstraintsForElems.

#define main lulesh_reference_main
#include "lulesh.cc"
#undef main

#include <omp.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>
#include <sys/time.h>

namespace {

FILE *g_dump = nullptr;

// Record: 48-byte name, 1-byte type ('d' float64, 'i' int32), 7 pad bytes,
// int64 count, then count raw little-endian values.
void put(const std::string &name, char type, const void *data, int64_t n)
{
   if (!g_dump) return;
   char hdr[56];
   memset(hdr, 0, sizeof(hdr));
   strncpy(hdr, name.c_str(), 47);
   hdr[48] = type;
   fwrite(hdr, 1, 56, g_dump);
   fwrite(&n, sizeof(n), 1, g_dump);
   size_t w = (type == 'd') ? 8 : 4;
   if (n > 0) fwrite(data, w, (size_t)n, g_dump);
}

void put_real(const std::string &name, Domain &d, Domain_member m, Index_t n)
{
   std::vector<double> buf(n);
   for (Index_t i = 0; i < n; ++i) buf[i] = (d.*m)(i);
   put(name, 'd', buf.data(), n);
}

void put_scalar(const std::string &name, double v) { put(name, 'd', &v, 1); }

void dump_setup(Domain &d)
{
   const Index_t ne = d.numElem(), nn = d.numNode();
   put("setup/nodelist", 'i', d.nodelist(0), 8 * (int64_t)ne);
   std::vector<Index_t> t(ne);
   for (Index_t i = 0; i < ne; ++i) t[i] = d.lxim(i);
   put("setup/lxim", 'i', t.data(), ne);
   for (Index_t i = 0; i < ne; ++i) t[i] = d.lxip(i);
   put("setup/lxip", 'i', t.data(), ne);
   for (Index_t i = 0; i < ne; ++i) t[i] = d.letam(i);
   put("setup/letam", 'i', t.data(), ne);
   for (Index_t i = 0; i < ne; ++i) t[i] = d.letap(i);
   put("setup/letap", 'i', t.data(), ne);
   for (Index_t i = 0; i < ne; ++i) t[i] = d.lzetam(i);
   put("setup/lzetam", 'i', t.data(), ne);
   for (Index_t i = 0; i < ne; ++i) t[i] = d.lzetap(i);
   put("setup/lzetap", 'i', t.data(), ne);
   for (Index_t i = 0; i < ne; ++i) t[i] = d.elemBC(i);
   put("setup/elemBC", 'i', t.data(), ne);
   put("setup/regNumList", 'i', d.regNumList(), ne);
   std::vector<Index_t> start(nn + 1);
   start[0] = 0;
   for (Index_t n = 0; n < nn; ++n) start[n + 1] = start[n] + d.nodeElemCount(n);
   put("setup/nodeElemStart", 'i', start.data(), nn + 1);
   put("setup/nodeElemCornerList", 'i', d.nodeElemCornerList(0), start[nn]);
   const Index_t nbc = (d.sizeX() + 1) * (d.sizeX() + 1);
   std::vector<Index_t> s(nbc);
   for (Index_t i = 0; i < nbc; ++i) s[i] = d.symmX(i);
   put("setup/symmX", 'i', s.data(), nbc);
   for (Index_t i = 0; i < nbc; ++i) s[i] = d.symmY(i);
   put("setup/symmY", 'i', s.data(), nbc);
   for (Index_t i = 0; i < nbc; ++i) s[i] = d.symmZ(i);
   put("setup/symmZ", 'i', s.data(), nbc);
   std::vector<Index_t> rs(d.numReg());
   for (Index_t r = 0; r < d.numReg(); ++r) rs[r] = d.regElemSize(r);
   put("setup/regElemSize", 'i', rs.data(), d.numReg());
   for (Index_t r = 0; r < d.numReg(); ++r)
      put("setup/regElemlist" + std::to_string(r), 'i', d.regElemlist(r), d.regElemSize(r));
}

const struct { const char *name; Domain_member m; bool node; } kFields[] = {
   {"x", &Domain::x, true},     {"y", &Domain::y, true},     {"z", &Domain::z, true},
   {"xd", &Domain::xd, true},   {"yd", &Domain::yd, true},   {"zd", &Domain::zd, true},
   {"xdd", &Domain::xdd, true}, {"ydd", &Domain::ydd, true}, {"zdd", &Domain::zdd, true},
   {"fx", &Domain::fx, true},   {"fy", &Domain::fy, true},   {"fz", &Domain::fz, true},
   {"nodalMass", &Domain::nodalMass, true},
   {"e", &Domain::e, false},    {"p", &Domain::p, false},    {"q", &Domain::q, false},
   {"ql", &Domain::ql, false},  {"qq", &Domain::qq, false},  {"v", &Domain::v, false},
   {"volo", &Domain::volo, false}, {"vnew", &Domain::vnew, false},
   {"delv", &Domain::delv, false}, {"vdov", &Domain::vdov, false},
   {"arealg", &Domain::arealg, false}, {"ss", &Domain::ss, false},
   {"elemMass", &Domain::elemMass, false},
};

void dump_state(const std::string &tag, Domain &d)
{
   for (const auto &f : kFields)
      put_real(tag + "/" + f.name, d, f.m, f.node ? d.numNode() : d.numElem());
   put_scalar(tag + "/time", d.time());
   put_scalar(tag + "/deltatime", d.deltatime());
   put_scalar(tag + "/dtcourant", d.dtcourant());
   put_scalar(tag + "/dthydro", d.dthydro());
   put_scalar(tag + "/cycle", (double)d.cycle());
}

// LagrangeElements (lulesh.cc:2432-2443) with CalcQForElems (lulesh.cc:1946-2010)
// opened up so the gradient temporaries can be written out before the
// reference frees them.  The call order is the reference's.
void elements_with_gradient_dump(Domain &d, const std::string &tag)
{
   const Index_t numElem = d.numElem();
   CalcLagrangeElements(d);
   Int_t allElem = numElem + 2 * d.sizeX() * d.sizeY() + 2 * d.sizeX() * d.sizeZ() +
                   2 * d.sizeY() * d.sizeZ();
   d.AllocateGradients(numElem, allElem);
   CalcMonotonicQGradientsForElems(d);
   put_real(tag + "/delx_xi", d, &Domain::delx_xi, numElem);
   put_real(tag + "/delx_eta", d, &Domain::delx_eta, numElem);
   put_real(tag + "/delx_zeta", d, &Domain::delx_zeta, numElem);
   put_real(tag + "/delv_xi", d, &Domain::delv_xi, numElem);
   put_real(tag + "/delv_eta", d, &Domain::delv_eta, numElem);
   put_real(tag + "/delv_zeta", d, &Domain::delv_zeta, numElem);
   CalcMonotonicQForElems(d);
   d.DeallocateGradients();
   for (Index_t i = 0; i < numElem; ++i) {
      if (d.q(i) > d.qstop()) {
         fprintf(stderr, "refdump: QStopError\n");
         exit(QStopError);
      }
   }
   ApplyMaterialPropertiesForElems(d);
   UpdateVolumesForElems(d, d.v_cut(), numElem);
}

} // namespace

int main(int argc, char **argv)
{
   int s = 30, nr = 11, bal = 1, cost = 1, its = 9999999, threads = 4;
   double emult = 1.0;
   const char *dumpfile = nullptr;
   std::set<int> dumpCycles;
   for (int i = 1; i < argc; ++i) {
      std::string a = argv[i];
      auto next = [&](void) -> const char * {
         if (i + 1 >= argc) { fprintf(stderr, "missing value for %s\n", a.c_str()); exit(2); }
         return argv[++i];
      };
      if (a == "-s") s = atoi(next());
      else if (a == "-r") nr = atoi(next());
      else if (a == "-b") bal = atoi(next());
      else if (a == "-c") cost = atoi(next());
      else if (a == "-i") its = atoi(next());
      else if (a == "-t") threads = atoi(next());
      else if (a == "-m") emult = atof(next());
      else if (a == "-d") dumpfile = next();
      else if (a == "-D") {
         std::string l = next();
         size_t p = 0;
         while (p < l.size()) {
            size_t q = l.find(',', p);
            if (q == std::string::npos) q = l.size();
            dumpCycles.insert(atoi(l.substr(p, q - p).c_str()));
            p = q + 1;
         }
      } else { fprintf(stderr, "unknown option %s\n", a.c_str()); return 2; }
   }
   if (threads < 2) {
      fprintf(stderr, "refdump: the gather path needs >= 2 OpenMP threads\n");
      return 2;
   }
   omp_set_num_threads(threads);

   Domain *dom = new Domain(1, 0, 0, 0, s, 1, nr, bal, cost);
   Domain &d = *dom;
   if (emult != 1.0) {
      // Same deposit as lulesh-init.cc:183-192, scaled by the multiplier.
      const Real_t ebase = Real_t(3.948746e+7);
      Real_t scale = s / Real_t(45.0);
      Real_t einit = ebase * scale * scale * scale * emult;
      d.e(0) = einit;
      d.deltatime() = (Real_t(.5) * cbrt(d.volo(0))) / sqrt(Real_t(2.0) * einit);
   }

   if (dumpfile) {
      g_dump = fopen(dumpfile, "wb");
      if (!g_dump) { perror(dumpfile); return 1; }
      dump_setup(d);
      dump_state("init", d);
   }

   timeval t0, t1;
   gettimeofday(&t0, nullptr);
   while (d.time() < d.stoptime() && d.cycle() < its) {
      TimeIncrement(d);
      const int c = d.cycle();
      if (g_dump && dumpCycles.count(c)) {
         const std::string tag = "c" + std::to_string(c);
         put_scalar(tag + "/deltatime_used", d.deltatime());
         LagrangeNodal(d);
         dump_state(tag + "/nodal", d);
         elements_with_gradient_dump(d, tag + "/grad");
         dump_state(tag + "/elements", d);
         CalcTimeConstraintsForElems(d);
         dump_state(tag + "/dt", d);
      } else {
         LagrangeLeapFrog(d);
      }
   }
   gettimeofday(&t1, nullptr);
   double elapsed = (double)(t1.tv_sec - t0.tv_sec) + (double)(t1.tv_usec - t0.tv_usec) / 1e6;

   // Plane-0 symmetry figures as lulesh-util.cc:197-212.
   Real_t maxAbs = 0, totAbs = 0, maxRel = 0;
   for (Index_t j = 0; j < s; ++j)
      for (Index_t k = j + 1; k < s; ++k) {
         Real_t ad = FABS(d.e(j * s + k) - d.e(k * s + j));
         totAbs += ad;
         if (maxAbs < ad) maxAbs = ad;
         Real_t rd = ad / d.e(k * s + j);
         if (maxRel < rd) maxRel = rd;
      }
   if (g_dump) {
      dump_state("final", d);
      fclose(g_dump);
   }
   printf("{\"s\": %d, \"r\": %d, \"b\": %d, \"c\": %d, \"threads\": %d, \"emult\": %.17g, "
          "\"cycles\": %d, \"e0\": \"%.17e\", \"maxabs\": \"%.17e\", \"totabs\": \"%.17e\", "
          "\"maxrel\": \"%.17e\", \"time\": \"%.17e\", \"deltatime\": \"%.17e\", "
          "\"elapsed_s\": %.6f, \"zones_per_us\": %.6f}\n",
          s, nr, bal, cost, threads, emult, d.cycle(), d.e(0), maxAbs, totAbs, maxRel, d.time(),
          d.deltatime(), elapsed,
          (double)d.numElem() * d.cycle() / (elapsed * 1e6));
   delete dom;
   return 0;
}
