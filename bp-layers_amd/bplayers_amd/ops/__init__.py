"""Autograd Functions and nn.Modules with the reference's ops/ module layout."""
