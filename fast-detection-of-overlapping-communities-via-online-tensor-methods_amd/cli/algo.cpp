// ./algo GPU_A|GPU_B [key=value ...] -- the reference program's command line (stpm.c:1068-1218)
// on the MI355X library: pre_proc, then GPU_Task on a second thread, then join.
// GPU_A / GPU_B select device 0 / 1 (stpm.c:1095-1118); key=value pairs override the fields of
// stpm_config (the stpm.h macros), e.g. NX=1000 NCOM=10 ALPHA_0=1.0 FOLDER_NAME=data/.
// Exit codes follow the reference: 1 for file errors, 2 for device / numerical errors.
#include <pthread.h>

#include <csignal>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "mmsb_stpm.h"

static void usage() {
  std::printf("usage: ./algo <option> [KEY=VALUE ...]\n");
  std::printf("choices for <option>:\n\t1) GPU_A\n\t2) GPU_B\n");
  std::printf("keys: NX NA NB NC NCOM ALPHA_0 LEARN_RATE THRESH PINV_TOL SGD_ITER_MIN SGD_ITER_MAX BINARY\n"
              "      SYNTHETIC FOLDER_NAME FILE_PREFIX MINIBATCH EPOCHS THETA RENORM_EVERY NS_ITERS NS_DEGREE\n"
              "      PRECISION POLISH_EPOCHS OVERSAMPLE POWER_ITERS M2_SCALING MEMBER_TAU SEED BA_FIRST_IS_ROW\n"
              "      FEOF_REPEAT VERBOSE USE_GRAPHS Y_MODE RENORM (ns | qr) CSR_HOST\n");
}

static int set_field(stpm_config& c, const std::string& key, const std::string& val) {
  auto I = [&](int32_t& f) { f = (int32_t)std::strtol(val.c_str(), nullptr, 10); };
  auto L = [&](int64_t& f) { f = (int64_t)std::strtoll(val.c_str(), nullptr, 10); };
  auto D = [&](double& f) { f = std::strtod(val.c_str(), nullptr); };
  auto S = [&](char* f, size_t n) { std::snprintf(f, n, "%s", val.c_str()); };
  if (key == "NX") I(c.nx);
  else if (key == "NA") I(c.na);
  else if (key == "NB") I(c.nb);
  else if (key == "NC") I(c.nc);
  else if (key == "NCOM") I(c.ncom);
  else if (key == "ALPHA_0") { D(c.alpha_0); S(c.alpha_0_str, sizeof(c.alpha_0_str)); }
  else if (key == "LEARN_RATE") { D(c.learn_rate); S(c.learn_rate_str, sizeof(c.learn_rate_str)); }
  else if (key == "THRESH") D(c.thresh);
  else if (key == "PINV_TOL") D(c.pinv_tol);
  else if (key == "SGD_ITER_MIN") L(c.sgd_iter_min);
  else if (key == "SGD_ITER_MAX") L(c.sgd_iter_max);
  else if (key == "BINARY") I(c.binary);
  else if (key == "EXPECTED") { I(c.binary); c.binary = !c.binary; }
  else if (key == "SYNTHETIC") I(c.synthetic);
  else if (key == "FOLDER_NAME") S(c.folder_name, sizeof(c.folder_name));
  else if (key == "FILE_PREFIX") S(c.file_prefix, sizeof(c.file_prefix));
  else if (key == "MINIBATCH") I(c.minibatch);
  else if (key == "EPOCHS") I(c.epochs);
  else if (key == "THETA") D(c.theta);
  else if (key == "RENORM_EVERY") I(c.renorm_every);
  else if (key == "NS_ITERS") I(c.ns_iters);
  else if (key == "NS_DEGREE") I(c.ns_degree);
  else if (key == "PRECISION") c.precision = (val == "bf16" || val == "1") ? MMSB_PREC_BF16 : MMSB_PREC_FP32;
  else if (key == "POLISH_EPOCHS") I(c.polish_epochs);
  else if (key == "OVERSAMPLE") I(c.oversample);
  else if (key == "POWER_ITERS") I(c.power_iters);
  else if (key == "M2_SCALING") c.m2_scaling = (val == "gpu" || val == "1") ? MMSB_M2_GPU : MMSB_M2_PAPER;
  else if (key == "MEMBER_TAU") D(c.member_tau);
  else if (key == "SEED") c.seed = std::strtoull(val.c_str(), nullptr, 10);
  else if (key == "BA_FIRST_IS_ROW") I(c.ba_first_is_row);
  else if (key == "FEOF_REPEAT") I(c.feof_repeat);
  else if (key == "VERBOSE") I(c.verbose);
  else if (key == "USE_GRAPHS") I(c.use_graphs);
  else if (key == "Y_MODE") I(c.y_mode);
  else if (key == "CSR_HOST") I(c.csr_host);
  else if (key == "RENORM") c.renorm_mode = (val == "qr" || val == "QR" || val == "1") ? MMSB_RENORM_QR : MMSB_RENORM_NS;
  else return -1;
  return 0;
}

// Ctrl-C plays the part of the reference's console 'e' (forceExit_flag = T_TRUE, stpm.c:1198): STGD
// stops at the next epoch boundary and the eigenvalues, post_proc and the output files follow. A
// second Ctrl-C terminates the process.
static void on_sigint(int) {
  stpm_force_exit();
  std::signal(SIGINT, SIG_DFL);
}

int main(int argc, char** argv) {
  if (argc < 2 || (std::strcmp(argv[1], "GPU_A") != 0 && std::strcmp(argv[1], "GPU_B") != 0)) {
    usage();
    return 1;
  }
  stpm_config cfg;
  stpm_config_default(&cfg);
  cfg.device = std::strcmp(argv[1], "GPU_B") == 0 ? 1 : 0;
  cfg.verbose = 1;
  for (int i = 2; i < argc; ++i) {
    const char* eq = std::strchr(argv[i], '=');
    if (!eq || set_field(cfg, std::string(argv[i], eq - argv[i]), std::string(eq + 1)) != 0) {
      std::fprintf(stderr, "unknown argument %s\n", argv[i]);
      usage();
      return 1;
    }
  }
  std::printf("%s: NX=%d NA=%d NB=%d NC=%d NCOM=%d ALPHA_0=%g LEARN_RATE=%g\n", mmsb_version(), cfg.nx, cfg.na, cfg.nb,
              cfg.nc, cfg.ncom, cfg.alpha_0, cfg.learn_rate);
  stpm_set_config(&cfg);
  int st = pre_proc();
  if (st != MMSB_OK) {
    std::fprintf(stderr, "pre_proc failed: %s\n", mmsb_last_error());
    return st == MMSB_ERR_IO ? 1 : 2;
  }
  pthread_t th;
  void* ret = nullptr;
  std::signal(SIGINT, on_sigint);
  if (pthread_create(&th, nullptr, GPU_Task, nullptr) != 0) return 2;
  pthread_join(th, &ret);
  if (ret) {
    const int code = *static_cast<int*>(ret);
    std::fprintf(stderr, "GPU_Task failed: %s\n", mmsb_last_error());
    return code == MMSB_ERR_IO ? 1 : 2;
  }
  std::printf("Task completed.\n");
  return 0;
}
