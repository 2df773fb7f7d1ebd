// lkc_cli.cpp -- the .in-driven command line: reads the key/value text of an
// .in file on stdin (as lakcluster_ver18 does, lakcluster_ver18.c:337-338) and runs
// the elimination on device 0, writing out_trace.txt / out_xdrop_{a,b}.txt, or with
// GLOBAL_TEST_TYPE= lakcluster_scorebox the scorebox replay and its .mda files, with
// GLOBAL_TEST_TYPE= dexcluster_driver the dcc elimination (several bins and Y/W controls
// included), or with GLOBAL_TEST_TYPE= bcc_sumscores / dcc_sumscores the reference's bcc / dcc
// self-tests on their synthetic problems.
#include "../../include/lkc.h"
#include <stdio.h>
#include <string>

int main(int argc, char** argv)
{
  std::string text;
  char buf[65536];
  size_t n;
  while ((n = fread(buf, 1, sizeof buf, stdin)) > 0) text.append(buf, n);
  const char* out_dir = argc > 1 ? argv[1] : nullptr;
  // GLOBAL_TEST_TYPE dispatch (lakcluster_ver18.c:337-345)
  size_t k = text.find("GLOBAL_TEST_TYPE");
  if (k != std::string::npos) {
    size_t e = text.find('=', k), s = text.find(';', k);
    std::string v = e != std::string::npos ? text.substr(e + 1, s == std::string::npos ? std::string::npos : s - e - 1) : "";
    v.erase(0, v.find_first_not_of(" \t\r\n"));
    v.erase(v.find_last_not_of(" \t\r\n") + 1);
    if (v == "dexcluster_driver") {
      if (lkc_dexcluster_driver(text.c_str(), out_dir)) { fprintf(stderr, "lkc: %s\n", lkc_last_error()); return 1; }
      printf("dexcluster_driver done\n");
      return 0;
    }
    if (v == "bcc_sumscores") {
      if (lkc_bcc_sumscores_test(text.c_str(), out_dir)) { fprintf(stderr, "lkc: %s\n", lkc_last_error()); return 1; }
      printf("bcc_sumscores done\n");
      return 0;
    }
    if (v == "dcc_sumscores") {
      if (lkc_dcc_sumscores_test(text.c_str(), out_dir)) { fprintf(stderr, "lkc: %s\n", lkc_last_error()); return 1; }
      printf("dcc_sumscores done\n");
      return 0;
    }
  }
  lkc_bcc* D = nullptr;
  int rc = lkc_bcc_load(text.c_str(), &D);
  if (rc) { fprintf(stderr, "lkc: %s\n", lkc_last_error()); return 1; }
  if (lkc_bcc_is_scorebox(D)) {
    rc = lkc_bcc_scorebox(D);
    if (!rc) rc = lkc_bcc_write_scorebox(D, out_dir);
    if (rc) { fprintf(stderr, "lkc: %s\n", lkc_bcc_error(D)); lkc_bcc_free(D); return 1; }
    int dims[4];
    lkc_bcc_get_scorebox(D, dims, NULL, NULL, NULL, NULL, NULL);
    printf("scorebox rows %d..0 in %d, cols %d..0 in %d\n", dims[0], dims[1], dims[2], dims[3]);
    lkc_bcc_free(D);
    return 0;
  }
  rc = lkc_bcc_run(D, -1);
  if (!rc) rc = lkc_bcc_write_outputs(D, out_dir);
  if (rc) { fprintf(stderr, "lkc: %s\n", lkc_bcc_error(D)); lkc_bcc_free(D); return 1; }
  double t[3];
  lkc_bcc_times(D, t);
  int shp[4];
  lkc_bcc_shape(D, shp);
  printf("rows %d controls %d cols %d covariates %d\n", shp[0], shp[1], shp[2], shp[3]);
  printf("iterations %d\nload %.3f s\ninit %.3f s\nloop %.3f s\nelim-run %.3f s\nbit-ops %lld\n",
         lkc_bcc_n_iter(D), t[0], t[1], t[2], t[1] + t[2], lkc_bcc_ops(D));
  lkc_bcc_free(D);
  return 0;
}
