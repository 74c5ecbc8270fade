// The serving scheduler (fishfry_amd/csrc/ff_server.cpp) under
// ThreadSanitizer, on the CPU stand-in device of fake_device.cpp: client
// threads submit host-array and field requests (FP64 and FP32 + refinement)
// and wait for them while another thread polls the statistics; two threads
// wait on one id; a server is destroyed with requests still queued.  Exits
// non-zero on a wrong result.
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>

#include "../../include/fishfry.h"

static std::atomic<int> g_fail{0};

// A field descriptor of extents n (the stand-in device ignores its values).
static void field_of(const int n[3], ff_field* f) {
  *f = ff_field{};
  for (int a = 0; a < 3; ++a) {
    f->n[a] = n[a];
    f->d[a] = 1.0;
  }
  f->scale = 1.0;
}

static void check(bool ok, const char* what) {
  if (!ok) {
    std::fprintf(stderr, "FAILED: %s\n", what);
    g_fail++;
  }
}

static void client(ff_server* s, int seed, int count) {
  std::mt19937 rng(seed);
  struct Pending {
    uint64_t id;
    std::vector<double> f, phi;
    bool field;
  };
  std::vector<Pending> reqs;
  std::vector<ff_field> fields(count);
  for (int k = 0; k < count; ++k) {
    ff_request r{};
    for (int a = 0; a < 3; ++a) r.n[a] = 4 + int(rng() % 20);
    r.precision = (rng() % 3 == 0) ? FF_FP32_REFINE : FF_FP64;
    r.refine_steps = r.precision == FF_FP32_REFINE ? int(rng() % 3) : 0;
    Pending p;
    p.field = rng() % 2 == 0;
    const size_t N = (size_t)r.n[0] * r.n[1] * r.n[2];
    if (p.field) {
      field_of(r.n, &fields[k]);
      r.field = &fields[k];
      r.want_norms = 1;
    } else {
      p.f.resize(N);
      p.phi.assign(N, 0.0);
      for (size_t i = 0; i < N; ++i) p.f[i] = double(rng() % 1000) * 1e-3;
      r.f_host = p.f.data();
    }
    reqs.push_back(std::move(p));
    if (!reqs.back().field) r.phi_host = reqs.back().phi.data();
    check(ff_server_submit(s, &r, &reqs.back().id) == FF_OK, "submit");
  }
  for (Pending& p : reqs) {
    ff_result res{};
    check(ff_server_wait(s, p.id, 30.0, &res) == FF_OK, "wait");
    check(res.status == FF_OK, "request status");
    if (p.field) {
      check(std::fabs(res.norms[0] - 0.5) < 1e-15 && std::fabs(res.norms[1] - 1.0) < 1e-15 &&
                std::fabs(res.norms[2] - 0.5) < 1e-15,
            "field norms");
    } else {
      bool ok = true;
      for (size_t i = 0; i < p.f.size(); ++i) ok = ok && p.phi[i] == 0.5 * p.f[i] + 1.0;
      check(ok, "host-array result");
    }
  }
}

int main() {
  ff_server_config cfg{size_t(64) << 20, 4, 4, 200.0, 0};
  ff_server* s = nullptr;
  check(ff_server_create(&cfg, &s) == FF_OK, "create");
  if (!s) return 1;
  std::atomic<bool> done{false};
  std::thread stats([&] {
    while (!done) {
      ff_serving_stats st{};
      check(ff_server_stats(s, &st) == FF_OK, "stats");
      std::this_thread::sleep_for(std::chrono::microseconds(200));
    }
  });
  std::vector<std::thread> clients;
  for (int c = 0; c < 4; ++c) clients.emplace_back(client, s, 100 + c, 30);
  // two threads waiting on one id: one takes the result, the other finds it gone
  {
    int n[3] = {8, 8, 8};
    ff_field f;
    field_of(n, &f);
    ff_request r{};
    r.n[0] = r.n[1] = r.n[2] = 8;
    r.field = &f;
    uint64_t id = 0;
    check(ff_server_submit(s, &r, &id) == FF_OK, "submit shared");
    std::atomic<int> ok{0}, gone{0};
    auto waiter = [&] {
      ff_result res{};
      const ff_status st = ff_server_wait(s, id, 30.0, &res);
      if (st == FF_OK) ok++;
      else if (st == FF_ERR_NOT_FOUND) gone++;
    };
    std::thread w1(waiter), w2(waiter);
    w1.join();
    w2.join();
    check(ok == 1 && gone == 1, "shared id: one result");
  }
  // a request larger than the budget is rejected at submission
  {
    ff_request r{};
    r.n[0] = r.n[1] = r.n[2] = 512;
    std::vector<double> f(1);
    r.f_host = f.data();
    uint64_t id = 0;
    check(ff_server_submit(s, &r, &id) == FF_ERR_REJECTED, "rejected");
  }
  for (std::thread& t : clients) t.join();
  done = true;
  stats.join();
  ff_serving_stats st{};
  check(ff_server_stats(s, &st) == FF_OK && st.completed == 4 * 30 + 1 && st.rejected == 1, "final stats");
  check(ff_server_destroy(s) == FF_OK, "destroy");
  // destroyed with requests still queued and in flight
  ff_server* s2 = nullptr;
  check(ff_server_create(&cfg, &s2) == FF_OK, "create 2");
  if (s2) {
    std::vector<double> f(16 * 16 * 16, 1.0);
    for (int k = 0; k < 20; ++k) {
      ff_request r{};
      r.n[0] = r.n[1] = r.n[2] = 16;
      r.f_host = f.data();
      uint64_t id = 0;
      check(ff_server_submit(s2, &r, &id) == FF_OK, "submit 2");
    }
    check(ff_server_destroy(s2) == FF_OK, "destroy 2");
  }
  // FF_SERVE_LARGE_ALONE: a flag the server does not know is refused; large
  // requests (FP32 + 8 steps at 170^3, about 3e8 point-passes each) complete
  // among small ones, each run with the device to itself
  {
    ff_server_config bad = cfg;
    bad.flags = 2;
    ff_server* s3 = nullptr;
    check(ff_server_create(&bad, &s3) == FF_ERR_INVALID, "unknown flag refused");
    ff_server_config alone{size_t(1) << 30, 4, 2, 100.0, FF_SERVE_LARGE_ALONE};
    check(ff_server_create(&alone, &s3) == FF_OK, "create 3");
    if (s3) {
      struct Job {
        std::vector<double> f, phi;
        uint64_t id = 0;
      };
      std::vector<Job> jobs(12);
      for (int k = 0; k < 12; ++k) {
        const bool large = k % 4 == 1;
        const int e = large ? 170 : 16;
        Job& j = jobs[k];
        j.f.assign(size_t(e) * e * e, 0.25 * k);
        j.phi.assign(j.f.size(), 0.0);
        ff_request r{};
        r.n[0] = r.n[1] = r.n[2] = e;
        r.precision = large ? FF_FP32_REFINE : FF_FP64;
        r.refine_steps = large ? 8 : 0;
        r.f_host = j.f.data();
        r.phi_host = j.phi.data();
        check(ff_server_submit(s3, &r, &j.id) == FF_OK, "submit 3");
      }
      for (Job& j : jobs) {
        ff_result res{};
        check(ff_server_wait(s3, j.id, 60.0, &res) == FF_OK && res.status == FF_OK, "wait 3");
        bool ok = true;
        for (size_t i = 0; i < j.f.size(); ++i) ok = ok && j.phi[i] == 0.5 * j.f[i] + 1.0;
        check(ok, "large-alone result");
      }
      check(ff_server_destroy(s3) == FF_OK, "destroy 3");
    }
  }
  if (g_fail) return 1;
  std::printf("server ok\n");
  return 0;
}
