// Output files of the DTFE command line (src/input_output.cc:213-298 and the
// writers of src/io/): one file per requested field, named
// <outputFilename>.<extension>, in the format userOptions.outputFileType
// selects: 101 raw binary, 111 text, 112 text with grid indices, 113 text
// with grid-cell centres, 114 text with redshift-cone positions, 100 binary
// with the 1024-byte density header.
#ifndef DTFE_DROPIN_INPUT_OUTPUT_H
#define DTFE_DROPIN_INPUT_OUTPUT_H

#include "quantities.h"
#include "user_options.h"

void writeOutputData(Quantities &uQuantities, Quantities &aQuantities, User_options const &userOptions);

#endif
