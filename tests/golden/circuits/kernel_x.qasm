OPENQASM 3.0;
include "stdgates.inc";
qubit[1] q;
x q[0];
