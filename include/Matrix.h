// Matrix.h -- dense row-major fp16 container (replaces Matrix.cuh:12-23).
#pragma once
#include <string>

#include "half_type.h"

class Matrix {
public:
    half *data = nullptr;
    int rows = 0;
    int cols = 0;
    int nonZeros = 0;

    // Zero-filled rows x cols matrix (Matrix.cu:6-12).
    Matrix(int rows, int cols);
    // Loads the ".mat" text format: "rows cols" then rows*cols whitespace-
    // separated values, row-major (Matrix.cu:14-32).  nonZeros counts values
    // that are nonzero as floats, before rounding to half.  Unlike the
    // reference (which redirects stdin with freopen and continues after a
    // failed open), this reads with an ifstream and throws
    // std::runtime_error on a missing or malformed file.
    explicit Matrix(const std::string &filename);
    ~Matrix();
    Matrix(const Matrix &) = delete;
    Matrix &operator=(const Matrix &) = delete;

    void print() const;
    // Writes the ".mat" format read by the constructor above.
    void save(const std::string &filename) const;
};
