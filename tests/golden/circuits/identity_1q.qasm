OPENQASM 3.0;
qubit[1] q;
