OPENQASM 3.0;
include "stdgates.inc";
qubit[1] q;
rx(pi/3) q[0];
