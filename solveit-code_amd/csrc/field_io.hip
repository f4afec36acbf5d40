// Field-data ingest on the host: the reference's Fields_*.csv and
// Triangles_*.csv files read as InterpolationScheme reads them
// (interpolation_scheme.cpp:66-172, returnSize auxiliary_functions.cpp:39-53),
// into arrays that shp_efield_setup uploads.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "shp_internal.h"

namespace shp {
namespace {

// returnSize: the number of fgets calls that return a line (a line longer
// than the 1024-byte buffer counts once per chunk); -1 when the file does not
// open.
long count_lines(const char* path) {
    FILE* fp = fopen(path, "r");
    if (!fp) return -1;
    char buf[1024];
    long n = 0;
    while (fgets(buf, sizeof(buf), fp)) n++;
    fclose(fp);
    return n;
}

// One fgets chunk per record, fields split on ',' as strtok splits them
// (empty fields are skipped), converted with the given parser; fields past
// `nf` are ignored and missing ones are left as they are.
template <typename T, typename F>
int read_records(const char* path, long nrec, int nf, T* const* out, F parse) {
    FILE* fp = fopen(path, "r");
    if (!fp) return -1;
    char buf[1024];
    for (long i = 0; i < nrec; i++) {
        if (!fgets(buf, sizeof(buf), fp)) break;
        int k = 0;
        char* save = nullptr;
        for (char* f = strtok_r(buf, ",", &save); f && k < nf; f = strtok_r(nullptr, ",", &save), k++)
            out[k][i] = parse(f);
    }
    fclose(fp);
    return 0;
}

}  // namespace
}  // namespace shp

using namespace shp;

extern "C" {

int shp_field_mesh_load_csv(const char* fields_path, const char* triangles_path, double r0, double E0,
                            shp_field_mesh* mesh) {
    SHP_REQUIRE(fields_path && triangles_path && mesh, "shp_field_mesh_load_csv: null argument");
    const long nt = count_lines(triangles_path);
    const long nd = count_lines(fields_path);
    SHP_REQUIRE(nd >= 0, "shp_field_mesh_load_csv: cannot open %s", fields_path);
    SHP_REQUIRE(nt >= 0, "shp_field_mesh_load_csv: cannot open %s", triangles_path);
    SHP_REQUIRE(nd < (1L << 30) && nt < (1L << 30), "shp_field_mesh_load_csv: %ld nodes, %ld triangles", nd, nt);
    double* xy = (double*)calloc(2 * (size_t)nd + 1, sizeof(double));
    double* er = (double*)calloc((size_t)nd + 1, sizeof(double));
    double* ez = (double*)calloc((size_t)nd + 1, sizeof(double));
    int* tri = (int*)calloc(3 * (size_t)nt + 1, sizeof(int));
    int* nbr = (int*)calloc(3 * (size_t)nt + 1, sizeof(int));
    if (!xy || !er || !ez || !tri || !nbr) {
        free(xy), free(er), free(ez), free(tri), free(nbr);
        set_error("shp_field_mesh_load_csv: out of host memory");
        return SHP_ERR_NOMEM;
    }
    // fields: x, y (units of r0), Er, Ez (units of E0); the reference stores x, y interleaved
    {
        std::vector<double> x(nd), y(nd);
        double* tmp[4] = {x.data(), y.data(), er, ez};
        read_records(fields_path, nd, 4, tmp, [](const char* f) { return atof(f); });
        for (long i = 0; i < nd; i++) {
            xy[2 * i] = x[i];
            xy[2 * i + 1] = y[i];
        }
    }
    // triangles: three nodes, then the three neighbours
    {
        std::vector<int> c[6];
        int* tmp[6];
        for (int k = 0; k < 6; k++) {
            c[k].assign(nt, 0);
            tmp[k] = c[k].data();
        }
        read_records(triangles_path, nt, 6, tmp, [](const char* f) { return atoi(f); });
        for (long i = 0; i < nt; i++)
            for (int k = 0; k < 3; k++) {
                tri[3 * i + k] = c[k][i];
                nbr[3 * i + k] = c[3 + k][i];
            }
    }
    mesh->nnodes = (int)nd;
    mesh->xy = xy;
    mesh->er = er;
    mesh->ez = ez;
    mesh->ntri = (int)nt;
    mesh->tri = tri;
    mesh->nbr = nbr;
    mesh->r0 = r0;
    mesh->E0 = E0;
    return SHP_OK;
}

void shp_field_mesh_free(shp_field_mesh* mesh) {
    if (!mesh) return;
    free((void*)mesh->xy);
    free((void*)mesh->er);
    free((void*)mesh->ez);
    free((void*)mesh->tri);
    free((void*)mesh->nbr);
    memset(mesh, 0, sizeof(*mesh));
}

}  // extern "C"
