// serve_book_test.cpp -- TEST INFRASTRUCTURE: the server's completion
// bookkeeping (hpttx::settle_requests) on the host, with no GPU and no HIP
// call: a slot whose event completed marks its requests done, records their
// latency and bytes; a slot whose launch or event failed marks them failed
// with the status and the reason, counts them as failed and records no
// latency.  Run by tests/test_capi.py.
#include <cstdio>
#include <cstdlib>

#include "serve.h"

using namespace hpttx;

#define EXPECT(c)                                                       \
  do {                                                                  \
    if (!(c)) {                                                         \
      std::fprintf(stderr, "serve_book_test:%d: %s\n", __LINE__, #c);   \
      std::exit(1);                                                     \
    }                                                                   \
  } while (0)

int main() {
  ServerStats st;
  std::vector<double> lat;
  std::deque<Request *> fin;
  Arena arena;   // no arena outputs here: nothing is returned to it
  Request a, b, c;
  a.bytes = 100;
  a.t_submit = 1.0;
  b.bytes = 50;
  b.t_submit = 1.5;
  c.bytes = 7;
  c.t_submit = 2.5;

  std::vector<Request *> done{&a, &b};
  EXPECT(settle_requests(done, kOk, "", 2.0, st, lat, fin, arena) == 150);
  EXPECT(done.empty());
  EXPECT(a.state == kDone && b.state == kDone && a.status == kOk);
  EXPECT(a.t_done == 2.0 && b.t_done == 2.0);
  EXPECT(st.completed == 2 && st.failed == 0 && st.bytes_done == 150);
  EXPECT(lat.size() == 2 && lat[0] == 1.0 && lat[1] == 0.5);
  EXPECT(st.last_done == 2.0);
  EXPECT(fin.size() == 2 && fin[0] == &a && fin[1] == &b);

  std::vector<Request *> failed{&c};
  EXPECT(settle_requests(failed, kEHip, "launch failed: unspecified launch failure", 3.0, st, lat,
                         fin, arena) == 7);
  EXPECT(failed.empty());
  EXPECT(c.state == kFailed && c.status == kEHip);
  EXPECT(c.error == "launch failed: unspecified launch failure");
  EXPECT(st.completed == 2 && st.failed == 1 && st.bytes_done == 150);
  EXPECT(lat.size() == 2);          // failures record no latency
  EXPECT(st.last_done == 2.0);      // nor a completion time
  EXPECT(fin.size() == 3 && fin[2] == &c);
  std::printf("SERVE BOOK OK\n");
  return 0;
}
