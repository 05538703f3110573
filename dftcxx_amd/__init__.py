"""MI355X-native grid-integration hot path of a Kohn-Sham LDA SCF iteration."""
