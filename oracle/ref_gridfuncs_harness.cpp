// Harness around the reference's own src/gridfuncs.cpp (compiled from the
// reference tree, never copied): defines the run-time globals that the
// reference's main.cpp:45-50 defines, and exposes the helpers through a C ABI
// for the oracle cross-checks.  Test infrastructure only.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "gridfuncs.h"

int D = 2;
int nGrids = 512;
double L = 32;
int nt = 1000;
int readxns = 1;
double c = 4.;

extern "C" {

int ref_iGridM(int d, int n, int iGrid, int iShift) { D = d; nGrids = n; return iGridMFunc(iGrid, iShift); }
int ref_iGridP(int d, int n, int iGrid, int iShift) { D = d; nGrids = n; return iGridPFunc(iGrid, iShift); }

void ref_xfunc(int d, int n, double l, int iGrid, double out[3])
{
    D = d; nGrids = n; L = l;
    std::vector<double> x = xFunc(iGrid, l);
    out[0] = x[0]; out[1] = x[1]; out[2] = x[2];
}

int ref_ijk(int d, int n, int iGrid, int out[3])
{
    D = d; nGrids = n;
    std::vector<int> v = iGridsFunc(iGrid);
    out[0] = v[0]; out[1] = v[1]; out[2] = v[2];
    return ijkTOiGrids(v);
}

// periodic images of one centre, as the text check_PBC writes
int ref_check_pbc(int d, const double xn[3], double shift, double Rn, char *buf, int cap)
{
    D = d;
    std::FILE *fp = std::tmpfile();
    if (!fp) return -1;
    check_PBC(std::vector<double>{xn[0], xn[1], xn[2]}, shift, Rn, fp);
    long len = std::ftell(fp);
    std::rewind(fp);
    int n = (int)(len < cap - 1 ? len : cap - 1);
    n = (int)std::fread(buf, 1, n, fp);
    buf[n] = 0;
    std::fclose(fp);
    return n;
}

}
