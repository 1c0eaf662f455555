// main.cc -- command-line driver: the reference's main loop (lulesh.cc:2650-2792)
// over the device path, with its flags and its end-of-run report
// (lulesh-util.cc:175-230).
//
// Flags as the reference (-q -i -s -r -b -c -p -v -h; -f accepted and
// ignored; -v writes the end-of-run dump as legacy VTK, host/viz.cc) plus:
//   -P fp64|fp64fma|mixed   device precision (default fp64)
//   -D              run TimeIncrement on the device (no per-cycle host sync)
//   -G              set the problem up on the device too (lgpu_create_sedov;
//                   implies -D; no host Domain, so no -v)
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>
#include <math.h>

#include <iomanip>
#include <iostream>

#include "host/domain.h"
#include "lulesh_host.h"

namespace {

struct Opts {
   int its = 9999999, nx = 30, numReg = 11, balance = 1, cost = 1;
   int showProg = 0, quiet = 0, precision = LGPU_FP64, deviceLoop = 0, viz = 0, deviceSetup = 0;
};

void usage(const char *exe)
{
   printf("Usage: %s [opts]\n", exe);
   printf(" -q              : quiet mode - suppress all stdout\n");
   printf(" -i <iterations> : number of cycles to run\n");
   printf(" -s <size>       : length of cube mesh along side\n");
   printf(" -r <numregions> : Number of distinct regions (def: 11)\n");
   printf(" -b <balance>    : Load balance between regions of a domain (def: 1)\n");
   printf(" -c <cost>       : Extra cost of more expensive regions (def: 1)\n");
   printf(" -p              : Print out progress\n");
   printf(" -v              : Output viz file (lulesh_plot_c<cycle>.vtk, legacy VTK)\n");
   printf(" -P fp64|fp64fma|mixed : device precision (def: fp64)\n");
   printf(" -D              : time step control on the device\n");
   printf(" -G              : problem set-up on the device as well (implies -D)\n");
   printf(" -h              : This message\n");
}

int int_arg(int argc, char **argv, int &i)
{
   if (i + 1 >= argc) {
      printf("Missing integer argument to %s\n", argv[i]);
      exit(-1);
   }
   char *end = nullptr;
   const long v = strtol(argv[i + 1], &end, 10);
   if (end == argv[i + 1] || (*end != '\0' && *end != ' ')) {
      printf("Parse Error on option %s integer value required after argument\n", argv[i]);
      exit(-1);
   }
   i += 2;
   return (int)v;
}

Opts parse(int argc, char **argv)
{
   Opts o;
   int i = 1;
   while (i < argc) {
      if (!strcmp(argv[i], "-i")) o.its = int_arg(argc, argv, i);
      else if (!strcmp(argv[i], "-s")) o.nx = int_arg(argc, argv, i);
      else if (!strcmp(argv[i], "-r")) o.numReg = int_arg(argc, argv, i);
      else if (!strcmp(argv[i], "-b")) o.balance = int_arg(argc, argv, i);
      else if (!strcmp(argv[i], "-c")) o.cost = int_arg(argc, argv, i);
      else if (!strcmp(argv[i], "-f")) (void)int_arg(argc, argv, i);
      else if (!strcmp(argv[i], "-p")) { o.showProg = 1; ++i; }
      else if (!strcmp(argv[i], "-q")) { o.quiet = 1; ++i; }
      else if (!strcmp(argv[i], "-v")) { o.viz = 1; ++i; }
      else if (!strcmp(argv[i], "-D")) { o.deviceLoop = 1; ++i; }
      else if (!strcmp(argv[i], "-G")) { o.deviceSetup = o.deviceLoop = 1; ++i; }
      else if (!strcmp(argv[i], "-P")) {
         if (i + 1 >= argc) { printf("Missing argument to -P\n"); exit(-1); }
         if (!strcmp(argv[i + 1], "fp64")) o.precision = LGPU_FP64;
         else if (!strcmp(argv[i + 1], "mixed")) o.precision = LGPU_MIXED;
         else if (!strcmp(argv[i + 1], "fp64fma")) o.precision = LGPU_FP64_FMA;
         else { printf("-P takes fp64, fp64fma or mixed\n"); exit(-1); }
         i += 2;
      } else if (!strcmp(argv[i], "-h")) {
         usage(argv[0]);
         exit(0);
      } else {
         printf("ERROR: Unknown command line argument: %s\n", argv[i]);
         usage(argv[0]);
         exit(-1);
      }
   }
   return o;
}


// VerifyAndWriteFinalOutput, lulesh-util.cc:175-230.
void report(const Opts &o, int cycle, const double *e, double elapsed)
{
   const long long nx8 = o.nx;
   const double grind = ((elapsed * 1e6) / cycle) / (double)(nx8 * nx8 * nx8);
   std::cout << "Run completed:\n";
   std::cout << "   Problem size        =  " << o.nx << "\n";
   std::cout << "   MPI tasks           =  1\n";
   std::cout << "   Iteration count     =  " << cycle << "\n";
   std::cout << "   Final Origin Energy =  " << std::scientific << std::setprecision(6)
             << std::setw(12) << e[0] << "\n";
   double maxAbs = 0, tot = 0, maxRel = 0;
   for (int j = 0; j < o.nx; ++j)
      for (int k = j + 1; k < o.nx; ++k) {
         const double ad = fabs(e[j * o.nx + k] - e[k * o.nx + j]);
         tot += ad;
         if (maxAbs < ad) maxAbs = ad;
         const double rd = ad / e[k * o.nx + j];
         if (maxRel < rd) maxRel = rd;
      }
   std::cout << "   Testing Plane 0 of Energy Array on rank 0:\n";
   std::cout << "        MaxAbsDiff   = " << std::setw(12) << maxAbs << "\n";
   std::cout << "        TotalAbsDiff = " << std::setw(12) << tot << "\n";
   std::cout << "        MaxRelDiff   = " << std::setw(12) << maxRel << "\n";
   std::cout.unsetf(std::ios_base::floatfield);
   std::cout << std::setprecision(2);
   std::cout << "\nElapsed time         = " << std::setw(10) << elapsed << " (s)\n";
   std::cout << std::setprecision(8);
   std::cout << "Grind time (us/z/c)  = " << std::setw(10) << grind << " (per dom)  ("
             << std::setw(10) << elapsed << " overall)\n";
   std::cout << "FOM                  = " << std::setw(10) << 1000.0 / grind << " (z/s)\n\n";
   printf("Full precision: cycles %d e0 %.17e maxabs %.17e totabs %.17e maxrel %.17e\n", cycle,
          e[0], maxAbs, tot, maxRel);
}

double seconds_since(const timeval &start)
{
   timeval end;
   gettimeofday(&end, nullptr);
   return (double)(end.tv_sec - start.tv_sec) + (double)(end.tv_usec - start.tv_usec) / 1e6;
}

// -G: the problem built by the device set-up kernels, stepped by lgpu_run.
int run_device_setup(const Opts &o)
{
   std::vector<Int_t> runReg, runLen;
   lhost::draw_region_runs(o.nx * o.nx * o.nx, o.numReg, o.balance, runReg, runLen);
   lgpu_sedov_desc desc{};
   desc.size = o.nx;
   desc.numReg = o.numReg;
   desc.balance = o.balance;
   desc.cost = o.cost;
   desc.emult = 1.0;
   desc.numRuns = (int32_t)runReg.size();
   desc.runReg = runReg.data();
   desc.runLen = runLen.data();
   desc.precision = o.precision;
   desc.flags = 0;
   lgpu_default_constants(&desc.k);
   lgpu_ctx *ctx = nullptr;
   lgpu_time_state ts{};
   const int st0 = lgpu_create_sedov(&desc, &ctx, &ts);
   if (st0) {
      fprintf(stderr, "device set-up failed (%d)\n", st0);
      return 1;
   }
   timeval start;
   gettimeofday(&start, nullptr);
   const int st = lgpu_run(ctx, &ts, o.its);
   const double elapsed = seconds_since(start);
   if (st) exit(st);
   std::vector<double> e((size_t)o.nx * o.nx * o.nx);
   if (lgpu_download(ctx, LGPU_E, e.data())) {
      fprintf(stderr, "device read-back failed\n");
      return 1;
   }
   if (!o.quiet) report(o, ts.cycle, e.data(), elapsed);
   lgpu_destroy(ctx);
   return 0;
}

} // namespace

int main(int argc, char **argv)
{
   const Opts o = parse(argc, argv);
   if (o.deviceSetup && o.viz) {
      printf("-G and -v cannot be combined (no host Domain to dump)\n");
      exit(-1);
   }
   if (!o.quiet) {
      std::cout << "Running problem size " << o.nx << "^3 per domain until completion\n";
      std::cout << "Num processors: 1 (one MI355X, " << (o.precision == LGPU_MIXED      ? "mixed"
                                                    : o.precision == LGPU_FP64_FMA ? "fp64fma"
                                                                                   : "fp64")
                << ")\n";
      std::cout << "Total number of elements: " << (long long)o.nx * o.nx * o.nx << " \n\n";
   }
   if (o.deviceSetup) return run_device_setup(o);
   Domain *dom = new Domain(1, 0, 0, 0, o.nx, 1, o.numReg, o.balance, o.cost);
   Domain &d = *dom;
   const int st0 = d.AttachDevice(o.precision);
   if (st0) {
      fprintf(stderr, "device set-up failed (%d): %s\n", st0, lgpu_last_error(d.device()));
      return 1;
   }

   timeval start;
   gettimeofday(&start, nullptr);
   if (o.deviceLoop) {
      lgpu_time_state ts = {d.time(), d.deltatime(), d.dtcourant(), d.dthydro(), d.stoptime(),
                            d.dtfixed(), d.dtmax(), d.deltatimemultlb(), d.deltatimemultub(),
                            d.cycle()};
      const int st = lgpu_run(d.device(), &ts, o.its);
      if (st) exit(st);
      d.time() = ts.time;
      d.deltatime() = ts.deltatime;
      d.dtcourant() = ts.dtcourant;
      d.dthydro() = ts.dthydro;
      d.cycle() = ts.cycle;
   } else {
      while (d.time() < d.stoptime() && d.cycle() < o.its) {
         TimeIncrement(d);
         LagrangeLeapFrog(d);
         if (o.showProg && !o.quiet) {
            std::cout << "cycle = " << d.cycle() << ", " << std::scientific
                      << "time = " << double(d.time()) << ", "
                      << "dt=" << double(d.deltatime()) << "\n";
            std::cout.unsetf(std::ios_base::floatfield);
         }
      }
   }
   const double elapsed = seconds_since(start);
   if (d.SyncFromDevice()) {
      fprintf(stderr, "device read-back failed\n");
      return 1;
   }
   if (o.viz) {
      // Final viz file, lulesh.cc:2776-2779.
      char name[64];
      snprintf(name, sizeof(name), "lulesh_plot_c%d.vtk", d.cycle());
      if (WriteVtk(d, name)) fprintf(stderr, "could not write %s\n", name);
   }
   if (!o.quiet) report(o, d.cycle(), &d.e(0), elapsed);
   delete dom;
   return 0;
}
