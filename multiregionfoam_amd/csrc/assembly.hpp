// Finite-volume assembly of the coupled temperature system on the device:
// the region matrices of transportTemperature / conductTemperature
// (setCoupledEqns, transportTemperature.C:129-150), the monolithic interface
// coefficients (calcThermalDiffusivity) and the ddt terms, from mesh
// geometry and the fields k, U, T resident in HBM, then placed into the
// solver's SpMV / DILU / interface arrays in place.
//
// Numbering is batch-wide over the systems that carry a model: cells,
// internal faces, boundary faces, interface-side faces and GGI pairs are
// concatenated system by system, region by region.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mrfg {

struct FvRegionParams {
    int transport;
    int pad;
    double rhoCp, rDeltaT;
};

struct FvDev {
    int nCells, nFaces, nSideFaces;
    const FvRegionParams *region;
    const int32_t *cellRegion, *faceRegion;
    // internal faces
    const int32_t *own, *nei;
    const double *w, *dc, *magSf, *Sf;
    // cells: fields, volumes, incident internal faces in ascending face
    // order (2*face + 1 where the cell is the neighbour) and boundary faces
    // in the reader's patch order
    const double *V, *k, *U, *T;
    const int32_t *cfOff, *cf, *cbOff;
    // boundary faces, grouped by owner cell (cbOff ranges)
    const int8_t *bT, *bU, *bKFixed;
    const double *bNDist, *bMagSf, *bSf, *bTb, *bUb, *bKb;
    const int32_t *bSide;
    // interface-side faces and their GGI pairs
    const int32_t *sCell, *sOff, *sPairCell;
    const double *sNDist, *sW, *sPairNDist;
    // scratch and results
    double *lapU, *aLo, *aUp;        // per internal face
    double *faceVal;                 // 2 per internal face: lower, upper
    double *sK, *sDelta;             // per side face: face conductivity, 1/|d|
    double *sKOwn, *sKNei;           // per side face: the wall value's two conductances
    double *ifCoeff;                 // per side face
    double *diag, *source, *ddtA, *diagBase, *srcBase;   // per cell
};

// Where the assembled values go, per batch position.
struct FvPlace {
    int nPos, K, G;
    int pos0;                        // positions [pos0, nPos) are placed
    const int32_t *cellOfPos;        // batch cell of the model, -1 (padding / no model)
    const int32_t *ellOff, *ellSrc;  // ELL columns per slice; source face value or -1
    const int32_t *fSrc, *bSrc;      // per DILU slot (nSlices*K*64)
    const unsigned long long *ifMask;
    const int32_t *ifBase, *ifRowFace, *ifSrc;
    const double *faceVal, *sideCoeff, *diag, *source, *ddtA, *diagBase, *srcBase;
    double *ellVal, *diagOut, *b, *sumA, *ifCoeff;
    double *fPack, *bPack, *prodPack;            // double views of the packed records
    double *ddtAOut, *diagBaseOut, *srcBaseOut;  // null when the batch is steady
};

hipError_t launchFvAssemble(const FvDev &m, hipStream_t st);
hipError_t launchFvPlace(const FvPlace &p, hipStream_t st);

}  // namespace mrfg
