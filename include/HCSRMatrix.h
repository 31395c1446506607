// HCSRMatrix.h -- hybrid split (replaces formats/HCSRMatrix.h:16-21): 16x16
// blocks whose density is >= threshold go to a BCSR part, the rest to a CSR
// part (formats/HCSRMatrix.cpp:10-36).
#pragma once
#include "BCSRMatrix.h"
#include "CSRMatrix.h"

class HCSRMatrix {
public:
    BCSRMatrix *bcsr = nullptr;
    CSRMatrix *csr = nullptr;
    HCSRMatrix(const Matrix &matrix, float threshold);
    ~HCSRMatrix();
    HCSRMatrix(const HCSRMatrix &) = delete;
    HCSRMatrix &operator=(const HCSRMatrix &) = delete;
};
