OPENQASM 3.0;
include "stdgates.inc";
qubit[1] q;
u3(pi/3, pi/4, pi/5) q[0];
