// Case files: electrodes.csv "x0, y0, z0, x1, y1, z1, radius" and nodes.csv
// "x, y, z", one record per line (src/electrode.c:45-101), and the examples'
// frequency list and result format (examples/harmonic_impedance.c:178-257).
// Host code only.
#include <cstdio>

#include "tags_common.h"

extern "C" {

int electrodes_file(const char file_name[], Electrode* electrodes, size_t num_electrodes)
{
    FILE* stream = std::fopen(file_name, "r");
    if (!stream) {
        std::printf("Cannot open file %s\n", file_name);
        return -10;
    }
    const int nvalues = 7;
    int got = nvalues, pop = 0;
    for (size_t i = 0; i < num_electrodes; i++) {
        double a[3], b[3], r;
        got = std::fscanf(stream, "%lf, %lf, %lf, %lf, %lf, %lf, %lf", &a[0], &a[1], &a[2],
                          &b[0], &b[1], &b[2], &r);
        if (got != nvalues) {
            std::printf("error reading line %i of file %s\n", (int)(i + 1), file_name);
            break;
        }
        pop = populate_electrode(&electrodes[i], a, b, r);
        if (pop != 0) {
            std::printf("Bad input: could not create electrode %i from file %s\n", (int)(i + 1),
                        file_name);
            break;
        }
    }
    std::fclose(stream);
    return pop == 0 ? got - nvalues : pop;
}

int nodes_file(const char file_name[], double* nodes, size_t num_nodes)
{
    FILE* stream = std::fopen(file_name, "r");
    if (!stream) {
        std::printf("Cannot open file %s\n", file_name);
        return -10;
    }
    int got = 3;
    for (size_t i = 0; i < num_nodes; i++) {
        got = std::fscanf(stream, "%lf, %lf, %lf", &nodes[3 * i], &nodes[3 * i + 1],
                          &nodes[3 * i + 2]);
        if (got != 3) {
            std::printf("error reading line %i of file %s\n", (int)(i + 1), file_name);
            break;
        }
    }
    std::fclose(stream);
    return got - 3;
}

long tags_count_lines(const char file_name[])
{
    FILE* f = std::fopen(file_name, "r");
    if (!f) return -1;
    long n = 1;  // as the examples count: newlines + 1
    for (int c = std::getc(f); c != EOF; c = std::getc(f))
        if (c == '\n') n++;
    std::fclose(f);
    return n;
}

int tags_frequencies_file(const char file_name[], double* freq, size_t n)
{
    FILE* f = std::fopen(file_name, "r");
    if (!f) {
        std::printf("Cannot open file %s\n", file_name);
        return -10;
    }
    size_t i = 0;
    for (; i < n; i++)
        if (std::fscanf(f, "%lf", freq + i) != 1) break;
    std::fclose(f);
    return (int)(n - i);
}

int tags_write_complex_file(const char file_name[], const TAGS_COMPLEX* z, size_t n)
{
    FILE* f = std::fopen(file_name, "w");
    if (!f) {
        std::printf("Cannot open file %s\n", file_name);
        return -10;
    }
    const double* d = (const double*)z;
    for (size_t i = 0; i < n; i++) std::fprintf(f, "%e%+e%s\n", d[2 * i], d[2 * i + 1], "im");
    std::fclose(f);
    return 0;
}

}  // extern "C"
