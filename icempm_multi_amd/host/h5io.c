/* h5io.c -- HDF5 layouts of the reference's SnapshotManager
 * (simulation/snapshotmanager.cpp:29-56, 125-302), with the HDF5 C API. */
#include "h5io.h"

#include <hdf5.h>
#include <stdio.h>
#include <string.h>

static char g_err[256];

const char *h5io_last_error(void) { return g_err; }

static int fail(const char *what)
{
    snprintf(g_err, sizeof g_err, "HDF5: %s", what);
    return -1;
}

static void quiet(void) { H5Eset_auto2(H5E_DEFAULT, NULL, NULL); }

static long long length_1d(hid_t ds)
{
    hid_t sp = H5Dget_space(ds);
    if (sp < 0) return -1;
    hsize_t n = 0;
    long long r = -1;
    if (H5Sget_simple_extent_ndims(sp) == 1 && H5Sget_simple_extent_dims(sp, &n, NULL) == 1) r = (long long)n;
    H5Sclose(sp);
    return r;
}

static int write_1d(hid_t file, const char *name, hid_t type, const void *data, hsize_t n)
{
    hid_t sp = H5Screate_simple(1, &n, NULL);
    if (sp < 0) return fail("dataspace");
    hid_t ds = H5Dcreate2(file, name, type, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    H5Sclose(sp);
    if (ds < 0) return fail(name);
    herr_t e = n ? H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) : 0;
    H5Dclose(ds);
    return e < 0 ? fail(name) : 0;
}

static int write_attr(hid_t obj, const char *name, hid_t type, const void *v)
{
    hid_t sp = H5Screate(H5S_SCALAR);
    if (sp < 0) return fail("scalar dataspace");
    hid_t a = H5Acreate2(obj, name, type, sp, H5P_DEFAULT, H5P_DEFAULT);
    H5Sclose(sp);
    if (a < 0) return fail(name);
    herr_t e = H5Awrite(a, type, v);
    H5Aclose(a);
    return e < 0 ? fail(name) : 0;
}

static int read_attr(hid_t obj, const char *name, hid_t type, void *v)
{
    hid_t a = H5Aopen(obj, name, H5P_DEFAULT);
    if (a < 0) return fail(name);
    herr_t e = H5Aread(a, type, v);
    H5Aclose(a);
    return e < 0 ? fail(name) : 0;
}

static int save_attrs(hid_t ds, const h5io_attrs *p)
{
    int e = 0;
    e |= write_attr(ds, "indenter_x", H5T_NATIVE_DOUBLE, &p->indenter_x);
    e |= write_attr(ds, "indenter_y", H5T_NATIVE_DOUBLE, &p->indenter_y);
    e |= write_attr(ds, "SimulationTime", H5T_NATIVE_DOUBLE, &p->SimulationTime);
    e |= write_attr(ds, "GridX", H5T_NATIVE_INT, &p->GridX);
    e |= write_attr(ds, "GridY", H5T_NATIVE_INT, &p->GridY);
    e |= write_attr(ds, "nPts", H5T_NATIVE_INT, &p->nPts);
    e |= write_attr(ds, "nPtsPitch", H5T_NATIVE_INT64, &p->nPtsPitch);
    e |= write_attr(ds, "UpdateEveryNthStep", H5T_NATIVE_INT, &p->UpdateEveryNthStep);
    e |= write_attr(ds, "n_indenter_subdivisions", H5T_NATIVE_INT, &p->n_indenter_subdivisions);
    e |= write_attr(ds, "cellsize", H5T_NATIVE_DOUBLE, &p->cellsize);
    e |= write_attr(ds, "IndDiameter", H5T_NATIVE_DOUBLE, &p->IndDiameter);
    e |= write_attr(ds, "InitialTimeStep", H5T_NATIVE_DOUBLE, &p->InitialTimeStep);
    e |= write_attr(ds, "SetupType", H5T_NATIVE_INT, &p->SetupType);
    e |= write_attr(ds, "Volume", H5T_NATIVE_DOUBLE, &p->Volume);
    return e ? -1 : 0;
}

int h5io_raw_count(const char *path, long long *n)
{
    quiet();
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail("cannot open raw points file");
    hid_t dg = H5Dopen2(f, "llGrainIDs", H5P_DEFAULT);
    int rc = 0;
    if (dg < 0) rc = fail("llGrainIDs");
    else {
        *n = length_1d(dg);
        if (*n < 0) rc = fail("llGrainIDs must be 1-D");
        H5Dclose(dg);
    }
    H5Fclose(f);
    return rc;
}

int h5io_read_raw(const char *path, double *x, double *y, uint64_t *grains, long long n, float *volume)
{
    quiet();
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail("cannot open raw points file");
    int rc = 0;
    hid_t dx = H5Dopen2(f, "x", H5P_DEFAULT), dy = H5Dopen2(f, "y", H5P_DEFAULT),
          dg = H5Dopen2(f, "llGrainIDs", H5P_DEFAULT);
    if (dx < 0 || dy < 0 || dg < 0) rc = fail("datasets x, y, llGrainIDs");
    else if (length_1d(dx) != n || length_1d(dy) != n || length_1d(dg) != n) rc = fail("dataset sizes differ");
    else if (H5Dread(dx, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, x) < 0) rc = fail("read x");
    else if (H5Dread(dy, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, y) < 0) rc = fail("read y");
    else if (H5Dread(dg, H5T_NATIVE_UINT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, grains) < 0) rc = fail("read llGrainIDs");
    else rc = read_attr(dg, "volume", H5T_NATIVE_FLOAT, volume);
    if (dx >= 0) H5Dclose(dx);
    if (dy >= 0) H5Dclose(dy);
    if (dg >= 0) H5Dclose(dg);
    H5Fclose(f);
    return rc;
}

int h5io_write_raw(const char *path, const double *x, const double *y, const uint64_t *grains, long long n,
                   float volume)
{
    quiet();
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail("cannot create file");
    int rc = write_1d(f, "x", H5T_NATIVE_DOUBLE, x, (hsize_t)n);
    if (!rc) rc = write_1d(f, "y", H5T_NATIVE_DOUBLE, y, (hsize_t)n);
    if (!rc) rc = write_1d(f, "llGrainIDs", H5T_NATIVE_UINT64, grains, (hsize_t)n);
    if (!rc) {
        hid_t dg = H5Dopen2(f, "llGrainIDs", H5P_DEFAULT);
        rc = dg < 0 ? fail("llGrainIDs") : write_attr(dg, "volume", H5T_NATIVE_FLOAT, &volume);
        if (dg >= 0) H5Dclose(dg);
    }
    H5Fclose(f);
    return rc;
}

int h5io_write_snapshot(const char *path, const void *params, size_t params_bytes, const double *points,
                        size_t n_doubles, unsigned size, const double *indenter, size_t n_ind,
                        const h5io_attrs *attrs)
{
    quiet();
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail("cannot create snapshot");
    int rc = write_1d(f, "Indenter_2D", H5T_NATIVE_DOUBLE, indenter, (hsize_t)n_ind);
    if (!rc) rc = write_1d(f, "Params", H5T_NATIVE_B8, params, (hsize_t)params_bytes);
    if (!rc) {
        hid_t dp = H5Dopen2(f, "Params", H5P_DEFAULT);
        rc = dp < 0 ? fail("Params") : save_attrs(dp, attrs);
        if (dp >= 0) H5Dclose(dp);
    }
    if (!rc) rc = write_1d(f, "Points", H5T_NATIVE_DOUBLE, points, (hsize_t)n_doubles);
    if (!rc) {
        hid_t ds = H5Dopen2(f, "Points", H5P_DEFAULT);
        rc = ds < 0 ? fail("Points") : write_attr(ds, "size", H5T_NATIVE_UINT, &size);
        if (ds >= 0) H5Dclose(ds);
    }
    H5Fclose(f);
    return rc;
}

int h5io_snapshot_sizes(const char *path, size_t *params_bytes, size_t *n_doubles)
{
    quiet();
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail("cannot open snapshot");
    int rc = 0;
    hid_t dp = H5Dopen2(f, "Params", H5P_DEFAULT), ds = H5Dopen2(f, "Points", H5P_DEFAULT);
    if (dp < 0 || ds < 0) rc = fail("datasets Params, Points");
    else {
        const long long a = length_1d(dp), b = length_1d(ds);
        if (a < 0 || b < 0) rc = fail("Params/Points must be 1-D");
        else { *params_bytes = (size_t)a; *n_doubles = (size_t)b; }
    }
    if (dp >= 0) H5Dclose(dp);
    if (ds >= 0) H5Dclose(ds);
    H5Fclose(f);
    return rc;
}

int h5io_read_snapshot(const char *path, void *params, size_t params_bytes, double *points, size_t n_doubles,
                       unsigned *size)
{
    size_t pb = 0, nd = 0;
    if (h5io_snapshot_sizes(path, &pb, &nd)) return -1;
    if (pb != params_bytes) return fail("SimParams size mismatch");
    if (nd != n_doubles) return fail("Points size mismatch");
    hid_t f = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
    if (f < 0) return fail("cannot open snapshot");
    int rc = 0;
    hid_t dp = H5Dopen2(f, "Params", H5P_DEFAULT), ds = H5Dopen2(f, "Points", H5P_DEFAULT);
    if (dp < 0 || ds < 0) rc = fail("datasets Params, Points");
    else if (H5Dread(dp, H5T_NATIVE_B8, H5S_ALL, H5S_ALL, H5P_DEFAULT, params) < 0) rc = fail("read Params");
    else if (H5Dread(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, points) < 0) rc = fail("read Points");
    else rc = read_attr(ds, "size", H5T_NATIVE_UINT, size);
    if (dp >= 0) H5Dclose(dp);
    if (ds >= 0) H5Dclose(ds);
    H5Fclose(f);
    return rc;
}

int h5io_write_pq(const char *path, const double *intact, size_t n_intact, const double *crushed, size_t n_crushed,
                  const h5io_attrs *attrs)
{
    quiet();
    hid_t f = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) return fail("cannot create pq file");
    int rc = 0;
    for (int k = 0; k < 2 && !rc; k++) {
        const char *name = k == 0 ? "PQ_intact" : "PQ_crushed";
        hsize_t dims[2] = {k == 0 ? n_intact : n_crushed, 2};
        hid_t sp = H5Screate_simple(2, dims, NULL);
        hid_t ds = sp < 0 ? -1 : H5Dcreate2(f, name, H5T_NATIVE_DOUBLE, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
        if (sp >= 0) H5Sclose(sp);
        if (ds < 0) { rc = fail(name); break; }
        if (dims[0] && H5Dwrite(ds, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, k == 0 ? intact : crushed) < 0)
            rc = fail(name);
        if (!rc && k == 0) rc = save_attrs(ds, attrs);
        H5Dclose(ds);
    }
    H5Fclose(f);
    return rc;
}
