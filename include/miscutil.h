// miscutil.h -- helpers of the reference's miscutil.h:9-19.
#pragma once
#include "Matrix.h"

void printMatrix(const float *M, unsigned int n);
bool checkMatrix(const float *A, const float *B, unsigned int n);
double rmse(const float *A, const float *B, unsigned int n);
float maxdiff(const float *A, const float *B, unsigned int n);
float avgrelerr(const float *A, const float *B, unsigned int n);
// Fraction of entries of the blockSizeRow x blockSizeCol block at (i, j)
// that are nonzero as halves (-0 counts as zero), miscutil.cpp:76-87.
float blockDensity(const Matrix &matrix, int i, int j, int blockSizeRow = 16,
                   int blockSizeCol = 16);
