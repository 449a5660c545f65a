// Command-line driver over libaccelerint_hip.so, the counterpart of the
// reference's GPU executable (generic/solver_main.cu:150-424): reads
// ign_data.bin, integrates every condition to the end time in global steps,
// optionally writes log/<solver_name>-log.bin, and prints the same timing
// lines ("Time:", "Time per step:", "TFinal:").
//
//   accelerint_main [num_IVPs] [device] [--end T] [--step dt] [--mech network.bin]
//                   [--ics file] [--log] [--log-end-only] [--mixed] [--fd-jacobian]
//                   [--atol a] [--rtol r]
//
// Without --mech the van der Pol example is integrated.  The end time and
// step (the reference's compile-time end_time / t_step) default to 1e-3 and
// 1e-6.  Without per-step logging all global steps run in one call (one
// persistent launch); with it, one intDriver call per global step, as the
// reference does.
#include <sys/stat.h>

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "accelerint.h"

static void usage_exit(const char* msg) {
  fprintf(stderr, "%s\n", msg);
  exit(1);
}

int main(int argc, char** argv) {
  int NUM = 1, device = 0, positional = 0;
  double end_time = 1e-3, t_step = 1e-6, atol = 1e-10, rtol = 1e-6;
  bool log_steps = false, log_end = false, mixed = false, fd = false;
  std::string mech, ics = "ign_data.bin";
  for (int a = 1; a < argc; ++a) {
    const std::string s = argv[a];
    auto next = [&]() -> const char* {
      if (a + 1 >= argc) usage_exit(("missing value after " + s).c_str());
      return argv[++a];
    };
    if (s == "--end") end_time = atof(next());
    else if (s == "--step") t_step = atof(next());
    else if (s == "--mech") mech = next();
    else if (s == "--ics") ics = next();
    else if (s == "--log") log_steps = true;
    else if (s == "--log-end-only") log_end = true;
    else if (s == "--mixed") mixed = true;
    else if (s == "--fd-jacobian") fd = true;
    else if (s == "--atol") atol = atof(next());
    else if (s == "--rtol") rtol = atof(next());
    else if (positional == 0) {
      if (sscanf(s.c_str(), "%i", &NUM) != 1 || NUM <= 0) {
        printf("Error: Problem size not in correct range\nProvide number greater than 0\n");
        exit(1);
      }
      ++positional;
    } else if (positional == 1) {
      if (sscanf(s.c_str(), "%i", &device) != 1 || device < 0) usage_exit("Error: GPU device number not in correct range");
      ++positional;
    } else {
      usage_exit("usage: accelerint_main [num_IVPs] [device] [--end T] [--step dt] [--mech file] [--ics file] "
                 "[--log] [--log-end-only] [--mixed] [--fd-jacobian] [--atol a] [--rtol r]");
    }
  }
  if (!(end_time > 0) || !(t_step > 0)) usage_exit("end time and step must be positive");

  if (!mech.empty()) {
    if (accelerInt_load_kinetics(mech.c_str()) != 0) usage_exit(accelerInt_last_error());
  } else if (accelerInt_select_mechanism(ACCELERINT_MECH_VAN_DER_POL) != 0) {
    usage_exit(accelerInt_last_error());
  }
  accelerInt_set_tolerances(atol, rtol);
  if (mixed) accelerInt_set_precision(ACCELERINT_PRECISION_MIXED);
  if (fd) accelerInt_set_jacobian(ACCELERINT_JACOBIAN_FINITE_DIFFERENCE);

  double* y_host = nullptr;
  double* var_host = nullptr;
  read_initial_conditions(ics.c_str(), NUM, &y_host, &var_host);
  accelerInt_initialize(NUM, device);
  printf("# ODEs: %d\n", NUM);

  std::string log_name;
  const bool logging = log_steps || log_end;
  if (logging) {
    struct stat info;
    if (stat("./log/", &info) != 0) {
      printf("Expecting 'log' subdirectory in current working directory. Please run"
             " mkdir log (or the equivalent) and run again.\n");
      exit(-1);
    }
    log_name = std::string("log/") + solver_name() + "-log.bin";
    if (accelerInt_write_log(log_name.c_str(), 0, NUM, 0.0, y_host) != 0) usage_exit(accelerInt_last_error());
  }

  const auto t0 = std::chrono::steady_clock::now();
  const double eps = 2.220446049250313e-16;
  int num_steps = 0;
  double t = 0.0;
  if (log_steps) {
    double t_next = fmin(end_time, t_step);
    while (t + eps < end_time) {
      ++num_steps;
      intDriver(NUM, t, t_next, var_host, y_host);
      t = t_next;
      t_next = fmin(end_time, (num_steps + 1) * t_step);
      if (accelerInt_write_log(log_name.c_str(), 1, NUM, t, y_host) != 0) usage_exit(accelerInt_last_error());
    }
  } else {
    accelerInt_integrate(NUM, 0.0, end_time, t_step, y_host, var_host);
    num_steps = (int)ceil(end_time / t_step - 1e-9);
    t = end_time;
    if (log_end && accelerInt_write_log(log_name.c_str(), 1, NUM, t, y_host) != 0)
      usage_exit(accelerInt_last_error());
  }
  const double runtime = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

  printf("Time: %.15e sec\n", runtime);
  printf("Time per step: %e (s)\t%.15e (s/thread)\n", runtime / num_steps, runtime / num_steps / NUM);
  printf("TFinal: %e\n", y_host[0]);
  accelerInt_cleanup();
  free(y_host);
  free(var_host);
  return 0;
}
