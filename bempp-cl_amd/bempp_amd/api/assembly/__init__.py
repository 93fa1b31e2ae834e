"""Assembly front end: assembler interfaces, boundary/potential operators, discrete operators."""
