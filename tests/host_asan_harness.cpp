// Host-side harness for the sanitizer test: exercises the C-ABI's host-only
// entry points (the case-file reader and the two text formats) on every
// file named on the command line and on edge-case arguments.  Built with
// AddressSanitizer and UndefinedBehaviorSanitizer; any report aborts.
#include <cstdio>
#include <cstring>
#include <initializer_list>

#include "../include/fen.h"

int main(int argc, char **argv)
{
    int ok = 0;
    for (int a = 1; a < argc; a++) {
        fen_case c;
        char err[512];
        if (fen_case_read(argv[a], &c, err, sizeof err) == FEN_OK) ok++;
        char tiny[3];
        (void)fen_case_read(argv[a], &c, tiny, sizeof tiny);   // message truncated to fit
        (void)fen_case_read(argv[a], &c, nullptr, 0);
    }
    (void)fen_case_read(nullptr, nullptr, nullptr, 0);
    char buf[256];
    const double vals[] = {0.0, -0.0, 1e-320, 5e-324, 1e300, -1.7976931348623157e308, 1.0 / 3.0, 9.9999995e99};
    for (double x : vals)
        for (double y : vals) {
            (void)fen_status_text(1234567, x, y, -x, y, buf, sizeof buf);
            (void)fen_status_text(-12345678, x, y, x, -y, buf, 8);   // too small: refused, truncated
        }
    const double x0[3] = {-1e-300, 123.456, 1e99};
    for (size_t len : {(size_t)1, (size_t)40, (size_t)1024})
        (void)fen_grid_json_text(12345678, -1, 0, x0, 1e-310, 2.0, 1e308, buf, len < sizeof buf ? len : sizeof buf);
    (void)fen_grid_json_text(1, 1, 1, nullptr, 1.0, 1.0, 1.0, nullptr, 0);
    std::printf("%d\n", ok);
    return 0;
}
