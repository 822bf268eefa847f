// ref_driver.cpp -- exposes the reference's host-only setup code through a
// C ABI so tests can compare against it.  TEST INFRASTRUCTURE ONLY: it is
// compiled against the reference's own sources (oracle/Makefile target
// `ref`) and never linked into the product.  Its outputs are written once
// to the committed fixture tests/golden/ref_angular.npz
// (tests/golden/make_ref_fixtures.py); the GPU tests read that fixture, so
// nothing built from the reference is loaded there.
#include <stdint.h>
#include <string.h>

#include "AngularData.h"   // AngularData.h:8-18
#include "LebedevQuad.h"   // LebedevQuad.h:4-12
#include "HemiQuad.h"      // HemiQuad.h:4-12
#include "Config.h"        // Config.h:4-48

double xE1(double x);      // E1.cpp:136-159

extern "C" {

// AngularData(maxk) (AngularData.cpp:9-68); arrays are the unpadded host ones.
int ref_angular(int maxk, double *omega, uint32_t *omega_pos, double *Ox,
                double *Oy, double *Oz) {
    AngularData ad(maxk);
    int slm = (int)ad.slm;
    memcpy(omega, ad.omega, sizeof(double) * 3 * slm * slm);
    memcpy(omega_pos, ad.omega_pos, sizeof(uint32_t) * 3 * slm * slm);
    memcpy(Ox, ad.Ox, sizeof(double) * slm * slm);
    memcpy(Oy, ad.Oy, sizeof(double) * slm * slm);
    memcpy(Oz, ad.Oz, sizeof(double) * slm * slm);
    return slm;
}

// LebedevQuad(mindegree) (LebedevQuad.cpp:6269-6277): returns the order;
// fills x,y,z,w when cap >= order.
int ref_lebedev(int mindegree, int cap, double *x, double *y, double *z,
                double *w) {
    LebedevQuad q(mindegree);
    if (q.order <= cap) {
        memcpy(x, q.x, sizeof(double) * q.order);
        memcpy(y, q.y, sizeof(double) * q.order);
        memcpy(z, q.z, sizeof(double) * q.order);
        memcpy(w, q.w, sizeof(double) * q.order);
    }
    return q.order;
}

// HemiQuad(mindegree) (HemiQuad.cpp:265-286)
int ref_hemiquad(int mindegree, int cap, double *x, double *y, double *z,
                 double *w) {
    HemiQuad q(mindegree);
    if (q.order <= cap) {
        memcpy(x, q.x, sizeof(double) * q.order);
        memcpy(y, q.y, sizeof(double) * q.order);
        memcpy(z, q.z, sizeof(double) * q.order);
        memcpy(w, q.w, sizeof(double) * q.order);
    }
    return q.order;
}

double ref_xE1(double x) { return xE1(x); }

// Config(cfgfile) (Config.cpp:8-131).  Returns 0 on success, -1 when the
// parser throws; ints = {maxk, dev, dump, prec, soltype},
// kappa/Ip get regions 1..nreg (the cap is the array size).
int ref_config(const char *path, int *ints, int cap, double *kappa, double *Ip,
               char *meshfn, char *outfn) {
    try {
        Config c(path);
        ints[0] = c.getMaxK();
        ints[1] = c.getDevice();
        ints[2] = c.doDump();
        ints[3] = c.usePrec();
        ints[4] = c.solType();
        for (int r = 1; r < cap; r++) {
            kappa[r] = c.getKappa(r);
            Ip[r] = c.getIp(r);
        }
        strncpy(meshfn, c.getMeshFilename(), 1023);
        strncpy(outfn, c.getOutFilename(), 1023);
        return 0;
    } catch (int) {
        return -1;
    }
}

}  // extern "C"
