// zg_cli.cpp -- `zorindex` command line (src/bin/zorindex.rs:16-355) over the
// C ABI: build, query and info. Queries run on the GPU.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/zindex_gpu.h"

static int usage() {
  fprintf(stderr,
          "usage: zorindex <build|query|info|fusion> [options]\n"
          "  build -f FOF -o OUT [-k K] [-m M] [-H HASHES] [-b BITS] [-s SEED]\n"
          "        [-i kmers|minimizers|modimizers|findere] [-d MODIMIZER_SAMPLING] [-z FINDERE_Z]\n"
          "        [-c] [-p SEED_PATTERN] [-r [-l STROBE_MIN] [-u STROBE_MAX]]\n"
          "        [-S SLOT_SCALE] [-L SEGMENT] [-X MAX_SUBINDEX]\n"
          "        [--merge-group-size N] [--drop-abandoned] [-t THREADS] [--gpu-peel]\n"
          "  query -i INDEX -q QUERY [-a MIN_RATIO] [-r] [-F tsv|compact] [-B BATCH]\n"
          "        [-M subindex|all] [--hbm-budget BYTES] [--no-recovery]\n"
          "  info -i INDEX\n"
          "  fusion -i INDEX INDEX... -o OUT\n");
  return 2;
}

static uint64_t parse_size(const std::string &v) {
  struct { const char *suffix; uint64_t scale; } units[] = {
      {"GiB", 1ULL << 30}, {"MiB", 1ULL << 20}, {"KiB", 1ULL << 10}, {"B", 1}};
  for (auto &u : units) {
    size_t n = strlen(u.suffix);
    if (v.size() > n && v.compare(v.size() - n, n, u.suffix) == 0)
      return strtoull(v.substr(0, v.size() - n).c_str(), nullptr, 10) * u.scale;
  }
  return strtoull(v.c_str(), nullptr, 10);
}

static int report(int rc) {
  if (rc) fprintf(stderr, "Error: %s\n", zg_last_error());
  return rc ? 1 : 0;
}

int main(int argc, char **argv) {
  if (argc < 2) return usage();
  std::string cmd = argv[1];
  std::vector<std::string> a(argv + 2, argv + argc);
  auto value = [&](size_t &i) -> std::string {
    if (i + 1 >= a.size()) {
      fprintf(stderr, "missing value for %s\n", a[i].c_str());
      exit(2);
    }
    return a[++i];
  };
  if (cmd == "build") {
    zg_build_config cfg;
    zg_build_config_default(&cfg);
    std::string fof, out;
    bool m_given = false;
    for (size_t i = 0; i < a.size(); i++) {
      const std::string &o = a[i];
      if (o == "-f" || o == "--fof") fof = value(i);
      else if (o == "-o" || o == "--output") out = value(i);
      else if (o == "-k" || o == "--k") cfg.k = (uint32_t)atoi(value(i).c_str());
      else if (o == "-m" || o == "--minimizer-size") { cfg.minimizer_size = (uint32_t)atoi(value(i).c_str()); m_given = true; }
      else if (o == "-H" || o == "--hashes") cfg.hashes = (uint32_t)atoi(value(i).c_str());
      else if (o == "-b" || o == "--fingerprint-bits") cfg.fingerprint_bits = (uint32_t)atoi(value(i).c_str());
      else if (o == "-s" || o == "--seed") cfg.seed = strtoull(value(i).c_str(), nullptr, 10);
      else if (o == "-S" || o == "--slot-scale") cfg.slot_scale = atof(value(i).c_str());
      else if (o == "-L" || o == "--segment-length") cfg.segment_length = (uint32_t)atoi(value(i).c_str());
      else if (o == "-X" || o == "--max-subindex-size") cfg.max_subindex_size = parse_size(value(i));
      else if (o == "-t" || o == "--threads") cfg.threads = (uint32_t)atoi(value(i).c_str());
      else if (o == "--merge-group-size") cfg.merge_group_size = (uint32_t)atoi(value(i).c_str());
      else if (o == "--drop-abandoned") cfg.drop_abandoned = 1;
      else if (o == "--gpu-peel") cfg.device = 1;
      else if (o == "-i" || o == "--index-mode") {
        std::string v = value(i);
        if (v == "kmers") cfg.index_mode = 0;
        else if (v == "minimizers") cfg.index_mode = 1;
        else if (v == "modimizers") cfg.index_mode = 2;
        else if (v == "findere") cfg.index_mode = 3;
        else { fprintf(stderr, "unsupported index mode %s\n", v.c_str()); return 2; }
      } else if (o == "-d" || o == "--modimizer-sampling") {
        cfg.modimizer_sampling = (uint32_t)atoi(value(i).c_str());
      } else if (o == "-z" || o == "--findere-z") {
        cfg.findere_z = (uint32_t)atoi(value(i).c_str());
      } else if (o == "-c" || o == "--homopolymer-compression") {
        cfg.homopolymer_compression = 1;
      } else if (o == "-r" || o == "--randstrobemers") {
        cfg.randstrobemers = 1;
      } else if (o == "-l" || o == "--strobe-min") {
        cfg.strobe_min = (uint32_t)atoi(value(i).c_str());
      } else if (o == "-u" || o == "--strobe-max") {
        cfg.strobe_max = (uint32_t)atoi(value(i).c_str());
      } else if (o == "-p" || o == "--seed-pattern") {
        std::string v = value(i);
        if (v.size() >= sizeof cfg.seed_pattern) {
          fprintf(stderr, "seed pattern width must be in 1..=31\n");
          return 2;
        }
        memcpy(cfg.seed_pattern, v.c_str(), v.size() + 1);
      } else return usage();
    }
    if (fof.empty() || out.empty()) return usage();
    if ((cfg.strobe_min || cfg.strobe_max) && !cfg.randstrobemers) {
      fprintf(stderr, "--strobe-min and --strobe-max require --randstrobemers\n");
      return 2;
    }
    if (cfg.index_mode == 1 && !m_given) {
      fprintf(stderr, "--minimizer-size is required for minimizers\n");
      return 2;
    }
    return report(zg_build_command(&cfg, fof.c_str(), out.c_str(), 2));
  }
  if (cmd == "query") {
    std::string index, query;
    double ratio = 0.0;
    bool records = false, recover = true;
    int format = 0;
    uint64_t batch = 4096, budget = 0;
    uint32_t memory = ZG_MEMORY_SUBINDEX;  // the reference's default query memory mode
    for (size_t i = 0; i < a.size(); i++) {
      const std::string &o = a[i];
      if (o == "-i" || o == "--index") index = value(i);
      else if (o == "-q" || o == "--query") query = value(i);
      else if (o == "-a" || o == "--min-ratio") ratio = atof(value(i).c_str());
      else if (o == "-r" || o == "--records") records = true;
      else if (o == "-F" || o == "--output-format") {
        std::string v = value(i);
        if (v == "tsv") format = 0;
        else if (v == "compact") format = 1;
        else return usage();
      } else if (o == "-B" || o == "--record-batch-size") batch = strtoull(value(i).c_str(), nullptr, 10);
      else if (o == "-M" || o == "--query-memory") {
        std::string v = value(i);
        if (v == "subindex") memory = ZG_MEMORY_SUBINDEX;
        else if (v == "all") memory = ZG_MEMORY_ALL;
        else return usage();
      } else if (o == "--hbm-budget") budget = strtoull(value(i).c_str(), nullptr, 10);
      else if (o == "-t" || o == "--threads") (void)value(i);
      else if (o == "--no-recovery") recover = false;
      else return usage();
    }
    if (index.empty() || query.empty()) return usage();
    if (records)
      return report(zg_query_records_command_ex(index.c_str(), query.c_str(), ratio, recover,
                                                batch, format, memory, budget, 1, 2));
    return report(zg_query_command_ex(index.c_str(), query.c_str(), ratio, recover, memory,
                                      budget, 1, 2));
  }
  if (cmd == "fusion") {
    std::vector<std::string> ins;
    std::string out;
    for (size_t i = 0; i < a.size(); i++) {
      if (a[i] == "-i" || a[i] == "--index") {
        while (i + 1 < a.size() && a[i + 1].rfind("-", 0) != 0) ins.push_back(a[++i]);
      } else if (a[i] == "-o" || a[i] == "--output") {
        out = value(i);
      } else {
        return usage();
      }
    }
    if (ins.size() < 2 || out.empty()) return usage();
    std::vector<const char *> ptrs;
    for (auto &x : ins) ptrs.push_back(x.c_str());
    return report(zg_fusion_command(ptrs.data(), (uint32_t)ptrs.size(), out.c_str(), 2));
  }
  if (cmd == "info") {
    std::string index;
    for (size_t i = 0; i < a.size(); i++) {
      if (a[i] == "-i" || a[i] == "--index") index = value(i);
      else return usage();
    }
    if (index.empty()) return usage();
    return report(zg_info_command(index.c_str(), 1));
  }
  return usage();
}
