0 1
1 2
# a comment ends the read, as in the reference
3 4
